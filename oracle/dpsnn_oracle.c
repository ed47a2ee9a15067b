/* dpsnn_oracle.c -- CPU restatement of DPSNN's LIFCA per-ms step.
 *
 * TEST INFRASTRUCTURE ONLY (see dpsnn_oracle.h).  Only tests, the smoke
 * check and the benchmark's CPU-baseline leg use it; the HIP engine never
 * does.  Numeric types follow the reference declaration by declaration so
 * that every rounding happens where the reference rounds:
 *   - ini values pass through float (initValue[] is float) before landing in
 *     double fields (DPSNN_LIFCAconnectome.cpp:313, DPSNN_parameters.h:177-194);
 *   - weights are int16 fixed point (DPSNN_dataStructDims.h:106);
 *   - neuron state, event times and bath values are double
 *     (DPSNN_neuron.h:106-111, DPSNN_parameters.h:153-156).
 * Build with -ffp-contract=off and no -march (the reference's g++ -O3 build
 * emits no FMA).
 *
 * Parity pins: the connectome and delivery order are checked against the
 * reference's own BACKWARD synapse dump; the step against the reference's
 * spike raster and ratesPerPop files (hash fixtures under tests/golden/);
 * ran3 against Knuth/NR known values; STDP tables against closed forms.
 */
#include "dpsnn_oracle.h"

#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAX_SUBPOP 27          /* DSD__subPopTot, DPSNN_dataStructDims.h:64 */
#define MAX_STENCIL 96         /* makeStencil_Max in the reference Makefile */
#define RING 80                /* DSD__maxD, DPSNN_dataStructDims.h:46 */
#define ANALOG_DEPTH 256       /* DPSNN_parameters.h:14 */
#define STDP_SPAN 128          /* DSD__maxSTDPspan_ms, DPSNN_neuron_sim.c:40 */
#define INT16_MIN_W (-32767)   /* DSD__INT16_MIN, DPSNN_dataStructDims.h:109 */

/* ------------------------------------------------------------------ */
/* ran3 (Knuth subtractive generator, Numerical Recipes) and deviates.  */
/* randdev.c:110-196, :477-502 (SetRandomSeed), :584-606 (NormDev).    */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t idum;        /* rand49_idum, randdev.c:33 (initial -77531) */
  int32_t inext, inextp, ma[56], iff;
  int32_t normSet;     /* NormDev's static Set, randdev.c:588 */
  double normII;       /* NormDev's static IIran, randdev.c:589 */
  uint64_t counter;
} ora_rng;

static double ran3(ora_rng *g) {
  int32_t mj, mk, i, ii, k;
  if (g->idum < 0 || g->iff == 0) {
    g->iff = 1;
    mj = 161803398 - (g->idum < 0 ? -g->idum : g->idum);
    mj = (mj < 0 ? -mj : mj);
    mj %= 1000000000;
    g->ma[55] = mj;
    mk = 1;
    for (i = 1; i <= 54; i++) {
      ii = (21 * i) % 55;
      g->ma[ii] = mk;
      mk = mj - mk;
      if (mk < 0) mk += 1000000000;
      mj = g->ma[ii];
    }
    for (k = 1; k <= 4; k++)
      for (i = 1; i <= 55; i++) {
        g->ma[i] -= g->ma[1 + (i + 30) % 55];
        if (g->ma[i] < 0) g->ma[i] += 1000000000;
      }
    g->inext = 0;
    g->inextp = 31;
    g->idum = 1;
  }
  if (++g->inext == 56) g->inext = 1;
  if (++g->inextp == 56) g->inextp = 1;
  mj = g->ma[g->inext] - g->ma[g->inextp];
  if (mj < 0) mj += 1000000000;
  g->ma[g->inext] = mj;
  g->counter++;
  return mj * (1.0 / (double)1000000000);
}

static void rng_seed(ora_rng *g, int32_t seed) {
  g->idum = -seed;
  (void)ran3(g);
}

static double norm_dev(ora_rng *g) {
  double fac, r, v1, v2;
  if (g->normSet == 0) {
    do {
      v1 = (double)(2.0 * ran3(g) - 1.0);
      v2 = (double)(2.0 * ran3(g) - 1.0);
      r = v1 * v1 + v2 * v2;
    } while (r >= 1.0 || r == 0.0);
    fac = (double)sqrt(-2.0 * log(r) / r);
    g->normII = v1 * fac;
    g->normSet = 1;
    return v2 * fac;
  }
  g->normSet = 0;
  return g->normII;
}

/* ------------------------------------------------------------------ */
/* erflib: roundr2i and the truncated-Gaussian LUT. erflib.c:41-181    */
/* ------------------------------------------------------------------ */
static uint32_t roundr2i(double r) {
  double i = floor(r);
  if (i + 0.5 > r) return (uint32_t)i;
  return (uint32_t)(i + 1.0);
}
static double erfcc(double x) {
  double t, z, ans;
  z = fabs(x);
  t = 1.0 / (1.0 + 0.5 * z);
  ans = t * exp(-z * z - 1.26551223 + t * (1.00002368 + t * (0.37409196 +
        t * (0.09678418 + t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 +
        t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
  return x >= 0.0 ? ans : 2.0 - ans;
}
static double normal_cum(double x, double mu, double sigma) {
  return 1 - 0.5 * erfcc((x - mu) / (sqrt(2.0) * sigma));
}
static double cut_normal_cum(double x, double mu, double sigma, double a, double b) {
  double n0 = normal_cum(a, mu, sigma);
  return (normal_cum(x, mu, sigma) - n0) / (normal_cum(b, mu, sigma) - n0);
}
static double find_probability(double p, double mu, double sigma, double a, double b) {
  double err = 1e-6, x, xmin, xmax;
  int it = 0;
  if (p <= 0.0) return a;
  if (p >= 1.0) return b;
  xmin = a; xmax = b;
  while ((xmax - xmin) > err && it < 256) {
    x = (xmax + xmin) * 0.5;
    if (cut_normal_cum(x, mu, sigma, a, b) < p) xmin = x; else xmax = x;
    it++;
  }
  return (xmax + xmin) * 0.5;
}
static void make_gaussian_lut(double *table, int n, double mu, double sigma,
                              double xmin, double xmax) {
  double p, dp, x0, x1;
  int i;
  if (sigma > 0.0) {
    dp = 1.0 / (double)n;
    x1 = xmin;
    for (p = dp, i = 0; p <= 1.0; p += dp, i++) {
      x0 = x1;
      x1 = find_probability(p, mu, sigma, xmin, xmax);
      table[i] = (x1 + x0) * 0.5;
    }
  } else {
    for (i = 0; i < n; i++) table[i] = mu;
  }
}

/* ------------------------------------------------------------------ */
/* Parameters (DPSNN_getParameters.c, DPSNN_parameters.h:72-146).      */
/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t count, offset;
  double DMin[MAX_SUBPOP], DMax[MAX_SUBPOP], J[MAX_SUBPOP], DJ[MAX_SUBPOP];
  double JExt, DJExt, NuExt, Tau, Theta, H, Tarp, AlphaC, TauC, gC;
} subpop_t; /* neuSubPopParamStruct, DPSNN_parameters.h:177-194 */

typedef struct {
  float STDPmultiplier, antiCausalCoeff, tauSTDP_ms, causalMax, antiCausalMax,
      tauDerivativeDecay_s, derivativeDrift;
} stdp_par_t; /* SongSTDP_paramStruct, DPSNN_parameters.h:49-58 */

typedef struct {
  uint32_t overallConnectivity, stencilDim, globH, CFX, CFY, CFT, npc, N,
      subPopNumber, tabNeuron, D, totalSimTime_ms, startPlasticity_ms,
      stopPlasticity_ms, moduloSec, startStatFiles_s, stopStatFiles_s,
      startRatesPerPop_s, stopRatesPerPop_s, ratesSampling, globalSeed,
      howManyOutputFiles, maxModDeltaX, maxModDeltaY, stencilX_Max,
      stencilY_Max, bathX_Max, bathY_Max, delayMin, delayMax;
  int spikes_writeDisable, ratesPerPop_writeDisable;
  float minSynWeight_f, f2w, w2f;
  stdp_par_t stdp;
} par_t;

struct ora_net {
  par_t p;
  subpop_t sp[MAX_SUBPOP];
  float probConnect[MAX_SUBPOP][MAX_SUBPOP][MAX_STENCIL][MAX_STENCIL];
  uint32_t bathEff[MAX_STENCIL][MAX_STENCIL];
  double bathLUT[MAX_SUBPOP][ANALOG_DEPTH];
  float causalTab[STDP_SPAN], antiCausalTab[STDP_SPAN];
  ora_rng rng;
  /* generation-ordered synapses */
  uint64_t S;
  uint32_t *g_pre, *g_post;
  uint8_t *g_delay;
  int16_t *g_w;
  /* backward (delivery-ordered) list, DPSNN_synapse.h:48-59 */
  uint32_t *b_pre, *b_post;
  uint8_t *b_delay, *b_kind;
  int16_t *b_w;
  float *b_td;
  int32_t *b_last;
  uint32_t *offsetByDelay; /* [N][RING], 0xFFFFFFFF = none */
  uint32_t *incoming_off, *incoming_idx; /* per-neuron incoming synapses (STDP) */
  /* neurons */
  uint8_t *subPop;
  double *v, *c, *Tr, *Te, *assignedExt, *invNuExt;
  uint32_t *maxExt;
  uint8_t *didItFire;
  int32_t *lastSpike_ms;
  /* per-step input list (addInputSpike) */
  uint32_t *in_cnt;
  uint64_t *in_off; /* per-neuron slice of the input pool */
  float *in_cur;
  double *in_frac;
  /* axonal ring, DPSNN_spike.c:193-277 */
  uint32_t *ring_ids[RING];
  double *ring_te[RING];
  uint32_t ring_cnt[RING];
  uint32_t ms;
  /* last step spikes */
  uint32_t nspk;
  uint64_t events; /* synaptic events delivered since create */
  /* rates per pop */
  uint32_t *firingsPerPop;
  /* output */
  char outdir[1024];
  FILE *fp_spikes, *fp_rates, *fp_probeA, *fp_probeB;
  int probe_enable;
  uint32_t probe_start_s, probe_stop_s;
  /* scratch */
  double *arr_t, *arr_v, *tmp_t, *tmp_v;
  uint32_t arr_cap;
  /* external recording */
  int rec_ext;
  uint32_t *ext_cnt;
  double *ext_t, *ext_v;
  uint32_t ext_n, ext_cap;
};

static void seterr(char *err, int errlen, const char *fmt, ...) {
  va_list ap;
  if (!err || errlen <= 0) return;
  va_start(ap, fmt);
  vsnprintf(err, errlen, fmt, ap);
  va_end(ap);
}

static int env_u32(const char *name, uint32_t *out, char *err, int errlen) {
  const char *s = getenv(name);
  if (!s) { seterr(err, errlen, "missing env var %s", name); return -1; }
  *out = (uint32_t)atoi(s);
  return 0;
}
static int env_bool(const char *name, int *out, char *err, int errlen) {
  const char *s = getenv(name);
  uint32_t v;
  if (!s) { seterr(err, errlen, "missing env var %s", name); return -1; }
  v = (uint32_t)atoi(s);
  if (v > 1) { seterr(err, errlen, "invalid %s", name); return -1; }
  *out = (int)v;
  return 0;
}

static FILE *open_in(const char *dir, const char *name) {
  char path[2048];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  return fopen(path, "r");
}

/* Tokenizes one line the way the reference readers do (strtok on " \t\n",
 * stop at a token starting with '#', atof into a float array). */
static int tokenize_floats(char *buf, float *vals, int maxv) {
  int n = 0;
  char *tok = strtok(buf, " \t\n");
  while (tok) {
    if (tok[0] == '#') break;
    if (n < maxv) vals[n] = (float)atof(tok);
    n++;
    tok = strtok(NULL, " \t\n");
  }
  return n;
}

/* readColumnFile DPSNN_getParameters.c:834-898 + initSubPopParamStruct
 * DPSNN_LIFCAconnectome.cpp:302-398. */
static int read_column(ora_net *n, const char *dir, char *err, int errlen) {
  char buf[200];
  float iv[12];
  FILE *fp = open_in(dir, "column.txt");
  uint32_t npc = 0, nsp = 0;
  if (!fp) { seterr(err, errlen, "cannot open column.txt"); return -1; }
  while (fgets(buf, sizeof buf, fp)) {
    int k, ntok;
    for (k = 0; k < 12; k++) iv[k] = 0.0f;
    iv[0] = (float)n->p.subPopNumber;
    ntok = tokenize_floats(buf, iv, 12);
    if (ntok > 0) {
      uint32_t s = (uint32_t)iv[0];
      if (ntok < 3) { seterr(err, errlen, "column.txt: short line"); fclose(fp); return -1; }
      if (s >= MAX_SUBPOP) { seterr(err, errlen, "column.txt: bad subPop"); fclose(fp); return -1; }
      npc += (uint32_t)iv[1];
      nsp++;
      n->sp[s].count = (uint32_t)iv[1];
      n->sp[s].offset = (s == 0) ? 0 : n->sp[s - 1].offset + n->sp[s - 1].count;
      n->sp[s].JExt = iv[2]; n->sp[s].DJExt = iv[3]; n->sp[s].NuExt = iv[4];
      n->sp[s].Tau = iv[5]; n->sp[s].Theta = iv[6]; n->sp[s].H = iv[7];
      n->sp[s].Tarp = iv[8]; n->sp[s].AlphaC = iv[9]; n->sp[s].TauC = iv[10];
      n->sp[s].gC = iv[11];
    }
  }
  fclose(fp);
  n->p.subPopNumber = nsp;
  n->p.npc = npc;
  if (npc == 0) { seterr(err, errlen, "neuronsPerCM <= 0"); return -1; }
  return 0;
}

/* initConnectivityParam DPSNN_LIFCAconnectome.cpp:16-131 and
 * initProbConnectMatrix :136-198 (explicit stencil, mirrored quadrants). */
static int read_connectivity(ora_net *n, const char *dir, char *err, int errlen) {
  static char buf[10 * MAX_STENCIL * MAX_STENCIL];
  static float iv[6 + MAX_STENCIL * MAX_STENCIL];
  static float spec[MAX_SUBPOP][MAX_SUBPOP][MAX_STENCIL][MAX_STENCIL];
  uint32_t mx = n->p.maxModDeltaX, my = n->p.maxModDeltaY, maxp, s, t, x, y;
  FILE *fp = open_in(dir, "connectivity.txt");
  if (!fp) { seterr(err, errlen, "cannot open connectivity.txt"); return -1; }
  maxp = n->p.overallConnectivity == 0 ? 6 + (n->p.stencilX_Max + 1) * (n->p.stencilY_Max + 1) : 7;
  memset(spec, 0, sizeof spec);
  while (fgets(buf, sizeof buf, fp)) {
    uint32_t j, idx;
    int ntok;
    iv[0] = (float)n->p.subPopNumber; iv[1] = (float)n->p.subPopNumber;
    for (j = 2; j < maxp; j++) iv[j] = 0.0f;
    ntok = tokenize_floats(buf, iv, 6 + MAX_STENCIL * MAX_STENCIL);
    if (ntok <= 0) continue;
    if ((uint32_t)(ntok - 6) != (mx + 1) * (my + 1)) {
      seterr(err, errlen, "connectivity.txt: wrong number of stencil entries");
      fclose(fp); return -1;
    }
    s = (uint32_t)iv[0]; t = (uint32_t)iv[1];
    if (s >= n->p.subPopNumber || t >= n->p.subPopNumber) {
      seterr(err, errlen, "connectivity.txt: bad subPop id"); fclose(fp); return -1;
    }
    idx = 2;
    n->sp[s].J[t] = iv[idx++]; n->sp[s].DJ[t] = iv[idx++];
    n->sp[s].DMin[t] = iv[idx++]; n->sp[s].DMax[t] = iv[idx++];
    for (y = 0; y <= my; y++)
      for (x = 0; x <= mx; x++) spec[s][t][x][y] = iv[idx++];
  }
  fclose(fp);
  memset(n->probConnect, 0, sizeof n->probConnect);
  for (s = 0; s < n->p.subPopNumber; s++)
    for (t = 0; t < n->p.subPopNumber; t++) {
      if (n->p.overallConnectivity == 0) {
        for (x = 0; x <= mx; x++)
          for (y = 0; y <= my; y++) {
            float v = spec[s][t][x][y];
            n->probConnect[s][t][mx + x][my + y] = v;
            n->probConnect[s][t][mx - x][my + y] = v;
            n->probConnect[s][t][mx - x][my - y] = v;
            n->probConnect[s][t][mx + x][my - y] = v;
          }
      } else {
        for (x = 0; x < n->p.stencilX_Max; x++)
          for (y = 0; y < n->p.stencilY_Max; y++)
            n->probConnect[s][t][x][y] = spec[s][t][0][0];
      }
    }
  return 0;
}

/* initBathEfficacyTemplate DPSNN_LIFCAconnectome.cpp:240-300. */
static int read_bath(ora_net *n, const char *dir, char *err, int errlen) {
  static char buf[100000];
  uint32_t x, y = 0;
  FILE *fp = open_in(dir, "bathEfficacy.txt");
  if (!fp) { seterr(err, errlen, "cannot open bathEfficacy.txt"); return -1; }
  memset(n->bathEff, 0, sizeof n->bathEff);
  while (fgets(buf, sizeof buf, fp)) {
    uint32_t ntok = 0;
    char *tok = strtok(buf, " \t\n");
    x = 0;
    while (tok) {
      if (tok[0] == '#') break;
      ntok++;
      if (x < MAX_STENCIL && y < MAX_STENCIL) n->bathEff[x][y] = (uint32_t)atoi(tok);
      x++;
      tok = strtok(NULL, " \t\n");
    }
    if (ntok > 0) {
      if (ntok != n->p.bathX_Max || ntok != n->p.bathY_Max) {
        seterr(err, errlen, "bathEfficacy.txt: size does not match stencilDim");
        fclose(fp); return -1;
      }
      y++;
    }
  }
  fclose(fp);
  return 0;
}

/* readPlasticityParametersFile DPSNN_getParameters.c:900-1053. */
static int read_plasticity(ora_net *n, const char *dir, char *err, int errlen) {
  char buf[200];
  float iv[20];
  int ntok = 0, j;
  FILE *fp = open_in(dir, "plasticity.txt");
  stdp_par_t *s = &n->p.stdp;
  if (!fp) {
    s->antiCausalCoeff = -1.0f; s->STDPmultiplier = 0.001f; s->tauSTDP_ms = 20.0f;
    s->causalMax = 0.1f; s->antiCausalMax = 0.12f; s->tauDerivativeDecay_s = 1.0f;
    s->derivativeDrift = 0.0f;
    return 0;
  }
  for (j = 1; j < 20; j++) iv[j] = 0.0f;
  while (fgets(buf, sizeof buf, fp)) ntok = tokenize_floats(buf, iv, 20);
  fclose(fp);
  if (ntok == 0 || ntok >= 20) { seterr(err, errlen, "plasticity.txt: bad columns"); return -1; }
  if (iv[0] != 1.0f) { seterr(err, errlen, "plasticity.txt: unknown algo"); return -1; }
  if (ntok != 8) { seterr(err, errlen, "plasticity.txt: expected 8 columns"); return -1; }
  s->STDPmultiplier = iv[2]; s->antiCausalCoeff = iv[1]; s->tauSTDP_ms = iv[3];
  s->causalMax = iv[4]; s->antiCausalMax = iv[5]; s->tauDerivativeDecay_s = iv[6];
  s->derivativeDrift = iv[7];
  return 0;
}

/* getParameters DPSNN_getParameters.c:23-831 (the subset the LIFCA step at
 * world size 1 reads). */
static int get_parameters(ora_net *n, const char *dir, char *err, int errlen) {
  par_t *p = &n->p;
  const char *s;
  uint32_t u;
  if (env_u32("env_overallConnectivity", &p->overallConnectivity, err, errlen)) return -1;
  if (p->overallConnectivity > 1) { seterr(err, errlen, "invalid env_overallConnectivity"); return -1; }
  if (env_u32("env_stencilDim", &p->stencilDim, err, errlen)) return -1;
  if (p->stencilDim == 0 || p->stencilDim > MAX_STENCIL) { seterr(err, errlen, "env_stencilDim out of range"); return -1; }
  if (p->overallConnectivity == 0) {
    p->maxModDeltaX = p->maxModDeltaY = p->stencilDim - 1;
    p->stencilX_Max = p->stencilY_Max = 2 * p->maxModDeltaX + 1;
  } else {
    p->maxModDeltaX = p->maxModDeltaY = 0;
    p->stencilX_Max = p->stencilY_Max = p->stencilDim;
  }
  if (p->stencilX_Max > MAX_STENCIL) { seterr(err, errlen, "stencil too large"); return -1; }
  p->bathX_Max = p->bathY_Max = p->stencilDim;
  if (env_u32("env_globH", &p->globH, err, errlen)) return -1;
  if (p->globH != 1) { seterr(err, errlen, "only env_globH=1 (one local network per device)"); return -1; }
  if (env_u32("env_synGen", &u, err, errlen)) return -1;
  if (u != 2) { seterr(err, errlen, "unsupported synGen %u", u); return -1; }
  if (read_column(n, dir, err, errlen)) return -1;
  if (env_u32("env_CFX", &p->CFX, err, errlen)) return -1;
  if (env_u32("env_CFY", &p->CFY, err, errlen)) return -1;
  p->CFT = p->CFX * p->CFY;
  if (p->CFT == 0 || p->CFT > 96 * 96 + 1) { seterr(err, errlen, "globCFT out of range"); return -1; }
  p->N = p->npc * p->CFT;
  s = getenv("env_minSynWeight_f");
  if (!s) { seterr(err, errlen, "missing env var env_minSynWeight_f"); return -1; }
  p->minSynWeight_f = (float)atof(s);
  if (p->minSynWeight_f >= 0 || p->minSynWeight_f < -6.0) { seterr(err, errlen, "env_minSynWeight_f out of range"); return -1; }
  p->f2w = INT16_MIN_W / p->minSynWeight_f;
  p->w2f = p->minSynWeight_f / INT16_MIN_W;
  /* tabNeuron, DPSNN_getParameters.c:431-457 (makeMaxCMFract = 1) */
  p->tabNeuron = p->N / p->CFT;
  if (env_u32("env_D", &p->D, err, errlen)) return -1;
  if (p->D == 0 || p->D > RING) { seterr(err, errlen, "env_D out of range"); return -1; }
  if (env_u32("env_totalSimTime_ms", &p->totalSimTime_ms, err, errlen)) return -1;
  if (env_u32("env_startPlasticity_ms", &p->startPlasticity_ms, err, errlen)) return -1;
  if (env_u32("env_stopPlasticity_ms", &p->stopPlasticity_ms, err, errlen)) return -1;
  if (env_u32("env_moduloSec", &p->moduloSec, err, errlen)) return -1;
  if (p->moduloSec == 0) { seterr(err, errlen, "env_moduloSec must be > 0"); return -1; }
  if (env_u32("env_startStatFiles_s", &p->startStatFiles_s, err, errlen)) return -1;
  if (env_u32("env_stopStatFiles_s", &p->stopStatFiles_s, err, errlen)) return -1;
  if (env_bool("env_spikes_writeDisable", &p->spikes_writeDisable, err, errlen)) return -1;
  {
    int dummy, probeDisable;
    if (env_bool("env_statSTDPevent_writeDisable", &dummy, err, errlen)) return -1;
    if (env_bool("env_statSynPeriodicProbe_writeDisable", &probeDisable, err, errlen)) return -1;
    n->probe_enable = !probeDisable;
  }
  if (env_bool("env_statRatesPerPop_writeDisable", &p->ratesPerPop_writeDisable, err, errlen)) return -1;
  if (env_u32("env_startRatesPerPop_s", &p->startRatesPerPop_s, err, errlen)) return -1;
  if (env_u32("env_stopRatesPerPop_s", &p->stopRatesPerPop_s, err, errlen)) return -1;
  if (env_u32("env_startSynPeriodicProbe_s", &n->probe_start_s, err, errlen)) return -1;
  if (env_u32("env_stopSynPeriodicProbe_s", &n->probe_stop_s, err, errlen)) return -1;
  if (env_u32("env_startSTDPevent_s", &u, err, errlen)) return -1;
  if (env_u32("env_stopSTDPevent_s", &u, err, errlen)) return -1;
  if (env_u32("env_ratesSampling", &p->ratesSampling, err, errlen)) return -1;
  if (p->ratesSampling == 0) { seterr(err, errlen, "env_ratesSampling must be > 0"); return -1; }
  s = getenv("env_globalSeed");
  if (!s) { seterr(err, errlen, "missing env var env_globalSeed"); return -1; }
  if (atoi(s) < 0) { seterr(err, errlen, "time-based seeds are not reproducible; give env_globalSeed >= 0"); return -1; }
  p->globalSeed = (uint32_t)atoi(s);
  s = getenv("env_howManyOutputFiles");
  p->howManyOutputFiles = s ? (uint32_t)atoi(s) : 0;
  if (p->howManyOutputFiles > 1) p->howManyOutputFiles = 0;
  if (read_plasticity(n, dir, err, errlen)) return -1;
  if (env_u32("env_collectivesAlgo", &u, err, errlen)) return -1;
  if (env_u32("env_pktLength", &u, err, errlen)) return -1;
  return 0;
}

/* describeConnectome DPSNN_LIFCAconnectome.cpp:400-487: sub-pop tables,
 * bath template, connectivity, delay range. */
static int describe_connectome(ora_net *n, const char *dir, char *err, int errlen) {
  uint32_t s, t;
  double dmin = 1e36, dmax = 0;
  if (read_bath(n, dir, err, errlen)) return -1;
  if (read_connectivity(n, dir, err, errlen)) return -1;
  for (s = 0; s < n->p.subPopNumber; s++)
    for (t = 0; t < n->p.subPopNumber; t++) {
      if (n->sp[s].DMax[t] > dmax) dmax = n->sp[s].DMax[t];
      if (n->sp[s].DMin[t] < dmin) dmin = n->sp[s].DMin[t];
    }
  n->p.delayMin = (uint32_t)dmin;
  n->p.delayMax = (uint32_t)dmax;
  return 0;
}

/* initLUT DPSNN_LIFCAcorticalModuleGen.cpp:347-356 (bath LUT only; the
 * synaptic-matrix LUTs feed nothing on this path). */
static void init_lut(ora_net *n) {
  uint32_t s;
  for (s = 0; s < n->p.subPopNumber; s++) {
    float J = (float)n->sp[s].JExt, DJ = (float)n->sp[s].DJExt;
    make_gaussian_lut(n->bathLUT[s], ANALOG_DEPTH, J, J * DJ, 0.0, 2.0 * J);
  }
}

static uint32_t sub_pop_of(const ora_net *n, uint32_t inCM) {
  uint32_t s;
  for (s = 0; s < n->p.subPopNumber; s++)
    if (inCM >= n->sp[s].offset && inCM < n->sp[s].count + n->sp[s].offset) return s;
  return 0xFFFFFFFFu;
}

/* getEmptyRandomSynapses DPSNN_LIFCAconnectome.cpp:1079-1133. */
static uint32_t empty_random_synapses(ora_net *n, float prob, uint32_t maxN) {
  double connectivity = (double)prob, r, P, C;
  uint32_t k = 1;
  r = ran3(&n->rng);
  P = connectivity;
  C = P;
  while ((C < r) && (k <= maxN)) {
    k++;
    P *= 1 - connectivity;
    C += P;
  }
  return k;
}

/* getRandomDelay_UNI DPSNN_LIFCAcorticalModuleGen.cpp:299-338. */
static uint32_t delay_uni(ora_net *n, uint32_t s, uint32_t t) {
  float DMin = (float)n->sp[s].DMin[t], DMax = (float)n->sp[s].DMax[t];
  float DelayMin = (float)n->p.delayMin, DelayMax = (float)n->p.delayMax;
  float DelayNum = (float)RING, DelayStep, outF;
  DelayStep = (DelayMax - DelayMin) / DelayNum;
  DelayMax -= DelayStep / 2;
  DelayMin += DelayStep / 2;
  DMax = roundr2i((DMax - DelayMin) / DelayStep) * DelayStep + DelayMin;
  DMin = roundr2i((DMin - DelayMin) / DelayStep) * DelayStep + DelayMin;
  outF = ((ran3(&n->rng) * (DMax - DMin) + DMin - DelayMin) / DelayStep);
  return roundr2i(outF * DelayStep);
}

/* simpleCM_invokeIndivSynGenerators DPSNN_LIFCAcorticalModuleGen.cpp:62-220
 * with generateTargetNeuList DPSNN_LIFCAconnectome.cpp:513-823 (Bernoulli
 * branch, non-periodic boundaries) and generateSourceNeuCExt :871-887. */
/* The external bath of neuron i (DPSNN_LIFCAcorticalModuleGen.cpp:203-211,
 * generateSourceNeuCExt: mirrored index into the bath template) and its
 * sub-population. */
static void neuron_bath(ora_net *n, uint32_t i) {
  par_t *p = &n->p;
  uint32_t npc = p->npc, cfy = (i / npc) / p->CFX, cfx = (i / npc) % p->CFX;
  uint32_t src = sub_pop_of(n, i % npc), bx = cfx, by = cfy;
  float nuExt = (float)n->sp[src].NuExt, cExt;
  n->subPop[i] = (uint8_t)src;
  if (bx >= p->CFX / 2) bx = -1 * (bx - p->CFX + 1);
  if (by >= p->CFY / 2) by = -1 * (by - p->CFY + 1);
  if (bx >= p->bathX_Max) bx = p->bathX_Max - 1;
  if (by >= p->bathY_Max) by = p->bathY_Max - 1;
  cExt = (float)n->bathEff[bx][by];
  if (cExt != 0) {
    n->invNuExt[i] = 1000.0 / (nuExt * cExt);
    n->maxExt[i] = (uint32_t)ceil(20 / n->invNuExt[i]);
  } else {
    n->invNuExt[i] = 0; n->maxExt[i] = 0;
  }
}

static int generate(ora_net *n, char *err, int errlen) {
  par_t *p = &n->p;
  uint32_t i, N = p->N, npc = p->npc;
  uint64_t cap = (uint64_t)N * 64 + 1024, S = 0;
  uint32_t *tlist = (uint32_t *)malloc(sizeof(uint32_t) * 65536);
  uint8_t *tsub = (uint8_t *)malloc(65536);
  n->g_pre = (uint32_t *)malloc(cap * 4); n->g_post = (uint32_t *)malloc(cap * 4);
  n->g_delay = (uint8_t *)malloc(cap); n->g_w = (int16_t *)malloc(cap * 2);
  for (i = 0; i < N; i++) {
    uint32_t cfy = (i / npc) / p->CFX, cfx = (i / npc) % p->CFX, inCM = i % npc;
    uint32_t src = sub_pop_of(n, inCM), ts, x, y, m = 0, j;
    float nuExt, cExt;
    uint32_t bx, by;
    n->subPop[i] = (uint8_t)src;
    if (i % p->tabNeuron == 0) rng_seed(&n->rng, (int32_t)(i / p->tabNeuron));
    for (ts = 0; ts < p->subPopNumber; ts++)
      for (x = 0; x < p->stencilX_Max; x++)
        for (y = 0; y < p->stencilY_Max; y++) {
          int32_t nx, ny;
          uint32_t tcx = p->CFX, tcy = p->CFY;
          float prob;
          if (p->overallConnectivity == 0) {
            nx = (int32_t)cfx + (int32_t)x - (int32_t)p->maxModDeltaX;
            ny = (int32_t)cfy + (int32_t)y - (int32_t)p->maxModDeltaY;
          } else {
            nx = (int32_t)x; ny = (int32_t)y;
          }
          if (nx >= 0 && nx < (int32_t)p->CFX) tcx = (uint32_t)nx;
          if (ny >= 0 && ny < (int32_t)p->CFY) tcy = (uint32_t)ny;
          if (!(tcx < p->CFX && tcy < p->CFY)) continue;
          prob = n->probConnect[src][ts][x][y];
          if (prob > 0) {
            uint32_t inT, last, avail;
            if (!(prob > 0.00001)) {
              seterr(err, errlen, "connection probabilities <= 1e-5 are not supported");
              free(tlist); free(tsub); return -1;
            }
            inT = n->sp[ts].offset - 1;
            last = n->sp[ts].offset + n->sp[ts].count;
            avail = last - inT;
            while ((inT += empty_random_synapses(n, prob, avail)) < last) {
              if (m >= 7500) { /* DSD__maxM (miniM), DPSNN_LIFCAcorticalModuleGen.cpp:113 */
                seterr(err, errlen, "neuron %u would get more than 7500 synapses", i);
                free(tlist); free(tsub); return -1;
              }
              tlist[m] = inT + tcx * npc + tcy * npc * p->CFX;
              tsub[m] = (uint8_t)sub_pop_of(n, inT);
              m++;
            }
          }
        }
    if (S + m > cap) {
      cap = (S + m) * 2;
      n->g_pre = (uint32_t *)realloc(n->g_pre, cap * 4); n->g_post = (uint32_t *)realloc(n->g_post, cap * 4);
      n->g_delay = (uint8_t *)realloc(n->g_delay, cap); n->g_w = (int16_t *)realloc(n->g_w, cap * 2);
    }
    for (j = 0; j < m; j++) {
      float weightF, mean, sigma;
      uint32_t d = delay_uni(n, src, tsub[j]);
      mean = (float)n->sp[src].J[tsub[j]];
      sigma = (float)(n->sp[src].DJ[tsub[j]] * mean);
      if (mean > 0.0)
        do { weightF = (float)(mean + norm_dev(&n->rng) * sigma); } while ((weightF < 0.0) || (weightF >= 2.0 * mean));
      else
        do { weightF = (float)(mean + norm_dev(&n->rng) * sigma); } while ((weightF <= 2.0 * mean) || (weightF > 0));
      if (weightF < p->minSynWeight_f || weightF > -p->minSynWeight_f) {
        seterr(err, errlen, "synaptic weight out of env_minSynWeight_f range");
        free(tlist); free(tsub); return -1;
      }
      if (d >= RING) { seterr(err, errlen, "delay out of range"); free(tlist); free(tsub); return -1; }
      n->g_pre[S] = i; n->g_post[S] = tlist[j]; n->g_delay[S] = (uint8_t)d;
      n->g_w[S] = (int16_t)(weightF * p->f2w);
      S++;
    }
    (void)nuExt; (void)cExt; (void)bx; (void)by;
    neuron_bath(n, i);
  }
  n->S = S;
  free(tlist); free(tsub);
  return 0;
}

/* inflateIntermToBackwardSynList DPSNN_localNet_init.c:631-740: stable
 * counting sort by delay, offsetByDelay = first index of each pre's run in
 * each delay bucket.  Also the per-neuron incoming lists (:212-237). */
static void build_layout(ora_net *n) {
  uint64_t S = n->S, i, off[RING + 1], cnt[RING];
  uint32_t N = n->p.N, d;
  memset(cnt, 0, sizeof cnt);
  for (i = 0; i < S; i++) cnt[n->g_delay[i]]++;
  off[0] = 0;
  for (d = 1; d <= RING; d++) off[d] = off[d - 1] + cnt[d - 1];
  n->b_pre = (uint32_t *)malloc(S * 4 + 4); n->b_post = (uint32_t *)malloc(S * 4 + 4);
  n->b_delay = (uint8_t *)malloc(S + 1); n->b_kind = (uint8_t *)malloc(S + 1);
  n->b_w = (int16_t *)malloc(S * 2 + 2); n->b_td = (float *)calloc(S + 1, 4);
  n->b_last = (int32_t *)calloc(S + 1, 4);
  {
    uint64_t pos[RING];
    memcpy(pos, off, sizeof pos);
    for (i = 0; i < S; i++) {
      uint64_t j = pos[n->g_delay[i]]++;
      n->b_pre[j] = n->g_pre[i]; n->b_post[j] = n->g_post[i];
      n->b_delay[j] = n->g_delay[i]; n->b_kind[j] = n->subPop[n->g_pre[i]];
      n->b_w[j] = n->g_w[i];
    }
  }
  n->offsetByDelay = (uint32_t *)malloc((uint64_t)N * RING * 4);
  memset(n->offsetByDelay, 0xFF, (uint64_t)N * RING * 4);
  for (d = 0; d < RING; d++) {
    uint64_t prev = (uint64_t)-1;
    for (i = off[d]; i < off[d + 1]; i++)
      if (n->b_pre[i] != prev) {
        prev = n->b_pre[i];
        n->offsetByDelay[(uint64_t)prev * RING + d] = (uint32_t)i;
      }
  }
  n->incoming_off = (uint32_t *)calloc(N + 1, 4);
  n->incoming_idx = (uint32_t *)malloc(S * 4 + 4);
  for (i = 0; i < S; i++) n->incoming_off[n->b_post[i] + 1]++;
  for (d = 0; d < N; d++) n->incoming_off[d + 1] += n->incoming_off[d];
  {
    uint32_t *fill = (uint32_t *)malloc((uint64_t)N * 4);
    memcpy(fill, n->incoming_off, (uint64_t)N * 4);
    for (i = 0; i < S; i++) n->incoming_idx[fill[n->b_post[i]]++] = (uint32_t)i;
    free(fill);
  }
}

/* initLongTermPlasticity DPSNN_neuron_sim.c:45-92. */
static void init_stdp_tables(ora_net *n) {
  int i;
  stdp_par_t *s = &n->p.stdp;
  n->causalTab[0] = s->causalMax * s->STDPmultiplier;
  n->causalTab[1] = n->causalTab[0];
  n->antiCausalTab[0] = s->antiCausalCoeff * s->antiCausalMax * s->STDPmultiplier;
  for (i = 1; i < STDP_SPAN; i++) n->causalTab[i] = (1 - 1 / s->tauSTDP_ms) * n->causalTab[i - 1];
  for (i = 1; i < STDP_SPAN; i++) n->antiCausalTab[i] = (1 - 1 / s->tauSTDP_ms) * n->antiCausalTab[i - 1];
}

static ora_net *create(const char *dir, uint64_t S, const uint32_t *pre, const uint32_t *post,
                       const uint8_t *delay, const int16_t *weight, char *err, int errlen) {
  ora_net *n = (ora_net *)calloc(1, sizeof(ora_net));
  uint32_t i, N;
  if (!n) { seterr(err, errlen, "out of memory"); return NULL; }
  n->rng.idum = -77531;
  if (get_parameters(n, dir, err, errlen) || describe_connectome(n, dir, err, errlen)) {
    free(n); return NULL;
  }
  if (n->p.stdp.tauSTDP_ms * 3 > STDP_SPAN || n->p.stdp.tauSTDP_ms <= 0 || n->p.stdp.STDPmultiplier < 0) {
    seterr(err, errlen, "invalid STDP parameters"); free(n); return NULL;
  }
  N = n->p.N;
  n->subPop = (uint8_t *)calloc(N, 1);
  n->v = (double *)calloc(N, 8); n->c = (double *)calloc(N, 8);
  n->Tr = (double *)calloc(N, 8); n->Te = (double *)calloc(N, 8);
  n->assignedExt = (double *)calloc(N, 8); n->invNuExt = (double *)calloc(N, 8);
  n->maxExt = (uint32_t *)calloc(N, 4); n->didItFire = (uint8_t *)calloc(N, 1);
  n->lastSpike_ms = (int32_t *)calloc(N, 4);
  /* localNet init seeds with env_globalSeed, DPSNN_localNet_init.c:119 */
  rng_seed(&n->rng, (int32_t)n->p.globalSeed);
  init_lut(n);
  if (!pre) {
    if (generate(n, err, errlen)) { ora_destroy(n); return NULL; }
  } else {
    /* a given backward list: its stable sort by delay is itself */
    n->S = S;
    n->g_pre = (uint32_t *)malloc(S * 4 + 4); n->g_post = (uint32_t *)malloc(S * 4 + 4);
    n->g_delay = (uint8_t *)malloc(S + 1); n->g_w = (int16_t *)malloc(S * 2 + 2);
    memcpy(n->g_pre, pre, S * 4); memcpy(n->g_post, post, S * 4);
    memcpy(n->g_delay, delay, S); memcpy(n->g_w, weight, S * 2);
    for (i = 0; i < N; i++) neuron_bath(n, i);
  }
  build_layout(n);
  for (i = 0; i < N; i++) {
    /* initNeuralKind DPSNN_neuron_init.c:99-115 */
    subpop_t *s = &n->sp[n->subPop[i]];
    n->v[i] = s->H; n->c[i] = 0.0; n->Te[i] = -1000; n->Tr[i] = 0.0;
    n->assignedExt[i] = 0.0;
    n->lastSpike_ms[i] = -100000; /* DPSNN_localNet_init.c:217 */
  }
  init_stdp_tables(n);
  for (i = 0; i < RING; i++) {
    n->ring_ids[i] = (uint32_t *)malloc((uint64_t)N * 4 + 4);
    n->ring_te[i] = (double *)malloc((uint64_t)N * 8 + 8);
    n->ring_cnt[i] = 0;
  }
  /* A neuron receives each incoming synapse at most once per step through
   * its own bucket and at most once more through a spill-over run, so twice
   * its in-degree bounds the per-step input list. */
  n->in_cnt = (uint32_t *)calloc(N, 4);
  n->in_off = (uint64_t *)malloc(((uint64_t)N + 1) * 8);
  n->in_off[0] = 0;
  for (i = 0; i < N; i++)
    n->in_off[i + 1] = n->in_off[i] + 2 * (uint64_t)(n->incoming_off[i + 1] - n->incoming_off[i]);
  n->in_cur = (float *)malloc(n->in_off[N] * 4 + 4);
  n->in_frac = (double *)malloc(n->in_off[N] * 8 + 8);
  n->firingsPerPop = (uint32_t *)calloc((uint64_t)n->p.CFT * n->p.subPopNumber, 4);
  /* the simulation stream continues from SetRandomSeed(loc_h = 0),
   * DPSNN_localNet_init.c:283 */
  rng_seed(&n->rng, 0);
  n->ms = 0;
  return n;
}

ora_net *ora_create(const char *dir, char *err, int errlen) {
  return create(dir, 0, NULL, NULL, NULL, NULL, err, errlen);
}

ora_net *ora_create_from_backward(const char *dir, uint64_t S, const uint32_t *pre, const uint32_t *post,
                                  const uint8_t *delay, const int16_t *weight, char *err, int errlen) {
  uint64_t i;
  for (i = 0; i < S; i++)
    if (delay[i] >= RING || (i && delay[i] < delay[i - 1])) {
      seterr(err, errlen, "the synapse list is not sorted by delay");
      return NULL;
    }
  return create(dir, S, pre, post, delay, weight, err, errlen);
}

/* Reads an engine checkpoint (layout documented with dpsnn_checkpoint in
 * include/dpsnn_hip.h) into the oracle's state: the neuron state, the
 * spikes of the last ring slots (the oracle's axonal ring), the fired flags
 * the next scan reads, the ran3 stream and the replayed drive's next event
 * times (or, for a Philox checkpoint, the device's next bath times, from
 * which this oracle continues with its own ran3 stream). */
static double poisson_lapse(ora_net *n, uint32_t i);

int ora_load_checkpoint(ora_net *n, const char *path, char *err, int errlen) {
  struct {
    char magic[8];
    uint32_t N, RS, NW, drive, stdp, t, nsub, ncols, writer, spikesOpen, ratesOpen, pad;
    uint64_t S, items;
  } h;
  struct {
    int32_t idum, inext, inextp, ma[56], iff, set, pad;
    double cached;
    uint64_t count;
  } rs;
  FILE *f = fopen(path, "rb");
  uint32_t N = n->p.N, i, m, RS;
  uint64_t k, *bits = NULL;
  double *skip = NULL, *hist = NULL;
  int ok = 0;
  if (!f) { seterr(err, errlen, "cannot open %s", path); return -1; }
  if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, "DPSNNCK1", 8) != 0) {
    seterr(err, errlen, "not a checkpoint"); fclose(f); return -1;
  }
  if (h.N != N || h.S != n->S) {
    seterr(err, errlen, "checkpoint of another network"); fclose(f); return -1;
  }
  RS = h.RS;
  skip = (double *)malloc((uint64_t)N * 8);
  bits = (uint64_t *)malloc((uint64_t)RS * h.NW * 8);
  hist = (double *)malloc((uint64_t)RS * N * 8);
  if (fread(n->v, 8, N, f) != N || fread(n->c, 8, N, f) != N || fread(n->Tr, 8, N, f) != N ||
      fread(n->Te, 8, N, f) != N || fread(n->assignedExt, 8, N, f) != N || fread(skip, 8, N, f) != N ||
      fread(bits, 8, (uint64_t)RS * h.NW, f) != (uint64_t)RS * h.NW ||
      fread(hist, 8, (uint64_t)RS * N, f) != (uint64_t)RS * N)
    goto out;
  {
    uint64_t counters[256];
    if (fread(counters, 8, 256, f) != 256) goto out;
  }
  if (h.stdp) {
    int16_t *w = (int16_t *)malloc(h.S * 2);
    float *td = (float *)malloc(h.S * 4);
    int32_t *last = (int32_t *)malloc(h.S * 4);
    int good = fread(n->lastSpike_ms, 4, N, f) == N && fread(w, 2, h.S, f) == h.S && fread(td, 4, h.S, f) == h.S &&
               fread(last, 4, h.S, f) == h.S;
    if (good)
      for (k = 0; k < h.S; k++) {  /* target-major -> backward order */
        uint32_t s = n->incoming_idx[k];
        n->b_w[s] = w[k]; n->b_td[s] = td[k]; n->b_last[s] = last[k];
      }
    free(w); free(td); free(last);
    if (!good) goto out;
  }
  if (fread(&rs, sizeof rs, 1, f) != 1) goto out;
  n->rng.idum = rs.idum; n->rng.inext = rs.inext; n->rng.inextp = rs.inextp;
  memcpy(n->rng.ma, rs.ma, sizeof rs.ma);
  n->rng.iff = rs.iff; n->rng.normSet = rs.set; n->rng.normII = rs.cached; n->rng.counter = rs.count;
  if (h.drive == 0 && fread(n->assignedExt, 8, N, f) != N) goto out;
  /* A checkpoint of the on-device Philox drive carries no replayed stream:
   * the oracle's own ran3 drive takes over from ms t, each neuron's first
   * external event one Poisson lapse after t.  The continuation is then
   * statistical, not bit-exact (used to time the oracle in the steady
   * state of a large network). */
  if (h.drive != 0)
    for (i = 0; i < N; i++) n->assignedExt[i] = (double)h.t + poisson_lapse(n, i);
  /* ring slots of ms t-1 .. t-RS+1 and the fired flags of ms t */
  n->ms = h.t;
  for (m = 0; m < RING; m++) n->ring_cnt[m] = 0;
  for (m = 1; m < RS && m <= h.t; m++) {
    uint32_t ms = h.t - m, slot = ms % RING, hs = ms & (RS - 1);
    for (i = 0; i < N; i++) {
      double te = hist[(uint64_t)hs * N + i];
      if (te >= 0) {
        n->ring_ids[slot][n->ring_cnt[slot]] = i;
        n->ring_te[slot][n->ring_cnt[slot]] = te;
        n->ring_cnt[slot]++;
      }
    }
  }
  for (i = 0; i < N; i++)
    n->didItFire[i] = (uint8_t)((bits[(uint64_t)(h.t & (RS - 1)) * h.NW + (i >> 6)] >> (i & 63)) & 1u);
  ok = 1;
out:
  if (!ok) seterr(err, errlen, "checkpoint file truncated");
  free(skip); free(bits); free(hist);
  fclose(f);
  return ok ? 0 : -1;
}

void ora_destroy(ora_net *n) {
  int i;
  if (!n) return;
  if (n->fp_spikes) fclose(n->fp_spikes);
  if (n->fp_rates) fclose(n->fp_rates);
  if (n->fp_probeA) fclose(n->fp_probeA);
  if (n->fp_probeB) fclose(n->fp_probeB);
  free(n->g_pre); free(n->g_post); free(n->g_delay); free(n->g_w);
  free(n->b_pre); free(n->b_post); free(n->b_delay); free(n->b_kind); free(n->b_w);
  free(n->b_td); free(n->b_last); free(n->offsetByDelay);
  free(n->incoming_off); free(n->incoming_idx);
  free(n->subPop); free(n->v); free(n->c); free(n->Tr); free(n->Te);
  free(n->assignedExt); free(n->invNuExt); free(n->maxExt); free(n->didItFire);
  free(n->lastSpike_ms);
  for (i = 0; i < RING; i++) { free(n->ring_ids[i]); free(n->ring_te[i]); }
  free(n->in_cnt); free(n->in_off); free(n->in_cur); free(n->in_frac); free(n->firingsPerPop);
  free(n->arr_t); free(n->arr_v); free(n->tmp_t); free(n->tmp_v);
  free(n->ext_cnt); free(n->ext_t); free(n->ext_v);
  free(n);
}

uint32_t ora_num_neurons(const ora_net *n) { return n->p.N; }
uint64_t ora_num_synapses(const ora_net *n) { return n->S; }
uint32_t ora_total_ms(const ora_net *n) { return n->p.totalSimTime_ms; }
uint32_t ora_now_ms(const ora_net *n) { return n->ms; }
uint64_t ora_synaptic_events(const ora_net *n) { return n->events; }
uint32_t ora_num_subpops(const ora_net *n) { return n->p.subPopNumber; }
uint32_t ora_neurons_per_cm(const ora_net *n) { return n->p.npc; }

void ora_get_backward(const ora_net *n, uint32_t *pre, uint32_t *post, uint8_t *delay,
                      uint8_t *kind, int16_t *weight, float *tderiv, int32_t *last_act) {
  uint64_t S = n->S;
  if (pre) memcpy(pre, n->b_pre, S * 4);
  if (post) memcpy(post, n->b_post, S * 4);
  if (delay) memcpy(delay, n->b_delay, S);
  if (kind) memcpy(kind, n->b_kind, S);
  if (weight) memcpy(weight, n->b_w, S * 2);
  if (tderiv) memcpy(tderiv, n->b_td, S * 4);
  if (last_act) memcpy(last_act, n->b_last, S * 4);
}
void ora_get_generated(const ora_net *n, uint32_t *pre, uint32_t *post, uint8_t *delay,
                       int16_t *weight) {
  uint64_t S = n->S;
  if (pre) memcpy(pre, n->g_pre, S * 4);
  if (post) memcpy(post, n->g_post, S * 4);
  if (delay) memcpy(delay, n->g_delay, S);
  if (weight) memcpy(weight, n->g_w, S * 2);
}
void ora_get_external(const ora_net *n, double *invNuExt, uint32_t *maxCount) {
  if (invNuExt) memcpy(invNuExt, n->invNuExt, (uint64_t)n->p.N * 8);
  if (maxCount) memcpy(maxCount, n->maxExt, (uint64_t)n->p.N * 4);
}
void ora_get_bath_lut(const ora_net *n, double *out) {
  uint32_t s;
  for (s = 0; s < n->p.subPopNumber; s++) memcpy(out + s * ANALOG_DEPTH, n->bathLUT[s], ANALOG_DEPTH * 8);
}

/* synapticListReport DPSNN_localNet_init.c:774-803 with the record format
 * of instrumentedSynapse::report DPSNN_synapse.h:63-91. */
int ora_write_synapse_dump(const ora_net *n, const char *path) {
  uint64_t i;
  FILE *fp = fopen(path, "w");
  if (!fp) return -1;
  for (i = 0; i < n->S; i++) {
    float w = (float)(n->b_w[i] * n->p.w2f);
    fprintf(fp, "%d %d %d %d %d %f ", (int)n->b_pre[i], (int)n->b_post[i],
            (int)(n->b_post[i] / n->p.N), (int)n->b_delay[i], (int)n->b_kind[i], w);
    fprintf(fp, "%f", n->b_td[i]);
    fprintf(fp, "\n");
  }
  fclose(fp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Output files (statBasicClass::openFile DPSNN_stat.c:55-107).         */
/* ------------------------------------------------------------------ */
void ora_set_output_dir(ora_net *n, const char *outdir) {
  snprintf(n->outdir, sizeof n->outdir, "%s", outdir ? outdir : "");
}
static FILE *open_stat(ora_net *n, const char *prefix, uint32_t sec) {
  char name[1400];
  par_t *p = &n->p;
  if (p->howManyOutputFiles == 1)
    snprintf(name, sizeof name, "%s/%s-sec%u-C%u-H%u-h%03u.dat", n->outdir, prefix, sec, p->CFT, p->globH, 0u);
  else if (sec == 0)
    snprintf(name, sizeof name, "%s/%s-start-C%u-H%u-h%u.dat", n->outdir, prefix, p->CFT, p->globH, 0u);
  else if (sec % 2 == 0)
    snprintf(name, sizeof name, "%s/%s-lasteven-C%u-H%u-h%u.dat", n->outdir, prefix, p->CFT, p->globH, 0u);
  else
    snprintf(name, sizeof name, "%s/%s-lastodd-C%u-H%u-h%u.dat", n->outdir, prefix, p->CFT, p->globH, 0u);
  return fopen(name, "w");
}
static int stat_window(const par_t *p, uint32_t sec) {
  return sec >= p->startStatFiles_s && sec <= p->stopStatFiles_s && sec % p->moduloSec == 0;
}

/* ------------------------------------------------------------------ */
/* The step.                                                            */
/* ------------------------------------------------------------------ */
static void ensure_arr(ora_net *n, uint32_t need) {
  if (need <= n->arr_cap) return;
  n->arr_cap = need * 2;
  n->arr_t = (double *)realloc(n->arr_t, n->arr_cap * 8);
  n->arr_v = (double *)realloc(n->arr_v, n->arr_cap * 8);
  n->tmp_t = (double *)realloc(n->tmp_t, n->arr_cap * 8);
  n->tmp_v = (double *)realloc(n->tmp_v, n->arr_cap * 8);
}

/* mergeFunction DPSNN_neuron_sim.c:391-421 (ties keep the left element). */
static void merge_fn(ora_net *n, double *At, double *Av, uint32_t p, uint32_t q, uint32_t r) {
  uint32_t i = p, j = q + 1, k = 0;
  double *Bt = n->tmp_t, *Bv = n->tmp_v;
  while (i <= q && j <= r) {
    if (At[i] <= At[j]) { Bt[k] = At[i]; Bv[k] = Av[i]; i++; }
    else { Bt[k] = At[j]; Bv[k] = Av[j]; j++; }
    k++;
  }
  while (i <= q) { Bt[k] = At[i]; Bv[k] = Av[i]; i++; k++; }
  while (j <= r) { Bt[k] = At[j]; Bv[k] = Av[j]; j++; k++; }
  for (k = p; k <= r; k++) { At[k] = Bt[k - p]; Av[k] = Bv[k - p]; }
}
/* mergeSort DPSNN_neuron_sim.c:423-434. */
static void merge_sort(ora_net *n, double *At, double *Av, uint32_t p, uint32_t r) {
  if (p < r) {
    uint32_t q = (p + r) / 2;
    merge_sort(n, At, Av, p, q);
    merge_sort(n, At, Av, q + 1, r);
    merge_fn(n, At, Av, p, q, r);
  }
}

/* getPoissonLapse DPSNN_neuron_sim.c:844-865. */
static double poisson_lapse(ora_net *n, uint32_t i) {
  double temp = (double)1.0 - ran3(&n->rng);
  if (temp < 0) temp = 0;
  return (-logf(temp) * n->invNuExt[i]);
}

/* LIFCADynamic DPSNN_neuron_sim.c:664-824 (second-order, non-Perseo). */
static void lifca(ora_net *n, uint32_t i, double currentTime, double I) {
  subpop_t *s = &n->sp[n->subPop[i]];
  double t = currentTime, deltaT, TFLES, rc, TrTe, dtHalf, rmHalf;
  double v = n->v[i], c = n->c[i];
  deltaT = t - n->Tr[i];
  TFLES = t - n->Te[i];
  if (TFLES > s->Tarp) {
    TrTe = n->Tr[i] - n->Te[i];
    if (TrTe < s->Tarp) {
      c *= exp(-(deltaT - TFLES + s->Tarp) / s->TauC);
      deltaT = TFLES - s->Tarp;
    }
    dtHalf = deltaT / 2;
    rmHalf = dtHalf / s->Tau;
    rc = deltaT / s->TauC;
    v = v - v * rmHalf - s->gC * c * dtHalf;
    v = v - v * rmHalf - s->gC * c * dtHalf;
    c = c * (1 - rc + rc * rc / 2);
    v += I;
    if (v >= s->Theta) {
      v = s->H;
      c += s->AlphaC;
      n->Te[i] = t;
      n->didItFire[i] = 1;
    }
  } else {
    rc = deltaT / s->TauC;
    c = c * (1 - rc + rc * rc / 2);
  }
  n->Tr[i] = t;
  n->v[i] = v;
  n->c[i] = c;
}

/* regularDynamicTimeStep DPSNN_neuron_sim.c:522-662. */
static void dynamics(ora_net *n, uint32_t i, uint32_t t) {
  double tF = (double)t, nextF = (double)(tF + 1.0);
  uint32_t ni = n->in_cnt[i], ne = 0, k, total;
  uint64_t base = n->in_off[i];
  const double *lut = n->bathLUT[n->subPop[i]];
  ensure_arr(n, ni + n->maxExt[i] + 1);
  for (k = 0; k < ni; k++) {
    n->arr_t[k] = tF + n->in_frac[base + k];
    n->arr_v[k] = n->in_cur[base + k];
  }
  if (ni) merge_sort(n, n->arr_t, n->arr_v, 0, ni - 1);
  if (n->invNuExt[i] > 0) {
    if (n->assignedExt[i] < nextF) {
      do {
        n->arr_t[ni + ne] = n->assignedExt[i];
        n->arr_v[ni + ne] = lut[(uint32_t)(ran3(&n->rng) * ANALOG_DEPTH)];
        ne++;
        n->assignedExt[i] += poisson_lapse(n, i);
      } while ((n->assignedExt[i] < nextF) && (ne < n->maxExt[i]));
      if (ne >= n->maxExt[i])
        if (n->assignedExt[i] < nextF) {
          do { n->assignedExt[i] += poisson_lapse(n, i); } while (n->assignedExt[i] < nextF);
        }
    }
  }
  if (n->rec_ext) {
    if (n->ext_n + ne > n->ext_cap) {
      n->ext_cap = (n->ext_n + ne) * 2 + 1024;
      n->ext_t = (double *)realloc(n->ext_t, n->ext_cap * 8);
      n->ext_v = (double *)realloc(n->ext_v, n->ext_cap * 8);
    }
    for (k = 0; k < ne; k++) {
      n->ext_t[n->ext_n + k] = n->arr_t[ni + k];
      n->ext_v[n->ext_n + k] = n->arr_v[ni + k];
    }
    n->ext_n += ne;
    n->ext_cnt[i] = ne;
  }
  total = ni + ne;
  if (ni && ne) merge_fn(n, n->arr_t, n->arr_v, 0, ni - 1, total - 1);
  for (k = 0; k < total; k++) lifca(n, i, n->arr_t[k], n->arr_v[k]);
}

/* LTP: causalSTDP_ms_ofBackwardSynapses DPSNN_neuron_sim.c:94-177. */
static void ltp(ora_net *n, uint32_t post, int32_t spike_ms) {
  uint32_t k;
  for (k = n->incoming_off[post]; k < n->incoming_off[post + 1]; k++) {
    uint32_t s = n->incoming_idx[k];
    int32_t last = n->b_last[s];
    float f = 0.0f;
    if (last != 0) {
      int dt = spike_ms - 1 - last;
      if (dt < STDP_SPAN) f = n->causalTab[dt];
    }
    n->b_td[s] = n->b_td[s] + f;
  }
}

/* LTD: synapseSetLastActiveTime_antiCausalSTDP DPSNN_neuron_sim.c:179-277. */
static void ltd(ora_net *n, uint32_t s, uint32_t post, uint32_t t) {
  int dt;
  float f;
  n->b_last[s] = (int32_t)t;
  dt = (int32_t)t - n->lastSpike_ms[post];
  f = (dt < STDP_SPAN) ? n->antiCausalTab[dt] : 0.0f;
  n->b_td[s] = n->b_td[s] + f;
}

/* statSynPeriodicProbeClass::write DPSNN_stat.c:429-451 (the synapse
 * record of instrumentedSynapse::report). */
static void probe_line(FILE *fp, const ora_net *n, uint64_t i) {
  float w = (float)(n->b_w[i] * n->p.w2f);
  fprintf(fp, "%d %d %d %d %d %f ", (int)n->b_pre[i], (int)n->b_post[i], (int)(n->b_post[i] / n->p.N),
          (int)n->b_delay[i], (int)n->b_kind[i], w);
  fprintf(fp, "%f", n->b_td[i]);
  fprintf(fp, "\n");
}

/* longTermPlasticity DPSNN_localNet_sim.c:312-414. */
static void commit_plasticity(ora_net *n) {
  uint64_t i;
  par_t *p = &n->p;
  for (i = 0; i < n->S; i++) {
    uint32_t src = n->b_kind[i], tgt = n->subPop[n->b_post[i]];
    float maxSyn = 2.0 * n->sp[src].J[tgt];
    float revMax = 2.0 * n->sp[tgt].J[src];
    float clamp = 0.0f, wF, tdF;
    if (n->fp_probeA) probe_line(n->fp_probeA, n, i);
    if (maxSyn > 0.0 && revMax > 0.0) {
      wF = (float)(n->b_w[i] * p->w2f);
      tdF = n->b_td[i];
      tdF += p->stdp.derivativeDrift;
      wF = wF + tdF;
      if (wF > maxSyn) { clamp = (wF - maxSyn); tdF = tdF - clamp; wF = maxSyn; }
      if (wF < 0.0) { clamp = wF; tdF = tdF - clamp; wF = 0.0; }
      n->b_w[i] = (int16_t)(wF * p->f2w);
      n->b_td[i] = tdF;
    }
    if (n->fp_probeB) probe_line(n->fp_probeB, n, i);
    if (p->stdp.tauDerivativeDecay_s == 1.0)
      n->b_td[i] = 0.0;
    else
      n->b_td[i] = (n->b_td[i] - p->stdp.derivativeDrift + clamp) * (1 - 1 / p->stdp.tauDerivativeDecay_s);
  }
}

/* completeTimeStep_ms DPSNN_localNet_sim.c:22-307 + the fire() shell
 * localNetProcess.c:102-158 + stopWatch DPSNN_stopWatch.c:29-44. */
int ora_fire(ora_net *n) {
  par_t *p = &n->p;
  uint32_t t = n->ms, N = p->N, i, sec = t / 1000, slot, d;
  int plastic;
  int new_s, x999;
  if (t == p->totalSimTime_ms) return 1;
  new_s = (t % 1000) == 0 && !((t % 1000) == 999);
  x999 = (t % 1000) == 999;
  plastic = (t >= p->startPlasticity_ms) && (t <= p->stopPlasticity_ms);
  if (new_s && n->outdir[0] && stat_window(p, sec)) {
    if (!p->spikes_writeDisable) n->fp_spikes = open_stat(n, "spikes", sec);
    if (!p->ratesPerPop_writeDisable && sec >= p->startRatesPerPop_s && sec <= p->stopRatesPerPop_s)
      n->fp_rates = open_stat(n, "ratesPerPop", sec);
    if (n->probe_enable && sec >= n->probe_start_s && sec <= n->probe_stop_s) {
      n->fp_probeA = open_stat(n, "synPeriodicProbe_A", sec);
      n->fp_probeB = open_stat(n, "synPeriodicProbe_B", sec);
    }
  }
  /* fired scan + compaction (:103-132, :620-664) and the axonal ring insert
   * (DPSNN_spike.c:209-239): slot t % 80 gets this step's ascending list. */
  slot = t % RING;
  n->ring_cnt[slot] = 0;
  for (i = 0; i < N; i++)
    if (n->didItFire[i]) {
      n->didItFire[i] = 0;
      n->lastSpike_ms[i] = (int32_t)t;
      if (plastic) ltp(n, i, (int32_t)t);
      n->ring_ids[slot][n->ring_cnt[slot]] = i;
      n->ring_te[slot][n->ring_cnt[slot]] = n->Te[i];
      n->ring_cnt[slot]++;
    }
  n->nspk = n->ring_cnt[slot];
  for (i = 0; i < N; i++) n->in_cnt[i] = 0;
  /* delivery (:837-999): delays D-1..0, spikes ascending, each pre's run
   * found through offsetByDelay and continued while pre matches
   * (findAndGetNextSyn DPSNN_synapse.c:48-136 keeps reading past its bucket
   * while the source id still matches). */
  for (d = p->D; d-- > 0;) {
    uint32_t sl = (t + RING - d) % RING, k;
    if (d > t) continue; /* never-written slots hold no spikes */
    for (k = 0; k < n->ring_cnt[sl]; k++) {
      uint32_t pre = n->ring_ids[sl][k];
      double te = n->ring_te[sl][k], ip;
      uint32_t s = n->offsetByDelay[(uint64_t)pre * RING + d];
      if (s == 0xFFFFFFFFu) continue;
      for (; s < n->S && n->b_pre[s] == pre; s++) {
        uint32_t post = n->b_post[s];
        float wF = ((float)(n->b_w[s])) * p->w2f;
        uint32_t c = n->in_cnt[post];
        if (n->in_off[post] + c >= n->in_off[post + 1]) { fprintf(stderr, "oracle: input overflow on neuron %u\n", post); abort(); }
        n->in_cur[n->in_off[post] + c] = wF;
        n->in_frac[n->in_off[post] + c] = modf(te, &ip);
        n->in_cnt[post] = c + 1;
        n->events++;
        if (plastic) ltd(n, s, post, t);
      }
    }
  }
  /* dynamics (:203-205) */
  if (n->rec_ext) {
    if (!n->ext_cnt) n->ext_cnt = (uint32_t *)calloc(N, 4);
    n->ext_n = 0;
  }
  for (i = 0; i < N; i++) dynamics(n, i, t);
  /* raster (:222-226, DPSNN_stat.c:275-296) */
  if (n->fp_spikes && stat_window(p, sec)) {
    uint32_t k;
    for (k = 0; k < n->ring_cnt[slot]; k++)
      fprintf(n->fp_spikes, "%d %.7g\n", (int)n->ring_ids[slot][k], n->ring_te[slot][k]);
  }
  /* per-population counts (:227-253) and the ratesPerPop writer
   * (DPSNN_stat.c:225-273) */
  {
    uint32_t k;
    for (k = 0; k < n->ring_cnt[slot]; k++) {
      uint32_t id = n->ring_ids[slot][k];
      uint32_t col = id / p->npc, inCol = id - col * p->npc, sp = 0;
      while (sp < p->subPopNumber) {
        if (inCol < n->sp[sp].count) break;
        inCol -= n->sp[sp].count;
        sp++;
      }
      n->firingsPerPop[col * p->subPopNumber + sp]++;
    }
    if (n->fp_rates && stat_window(p, sec) && sec >= p->startRatesPerPop_s &&
        sec <= p->stopRatesPerPop_s && (t % p->ratesSampling) == 0) {
      uint32_t tot, sp;
      fprintf(n->fp_rates, "%d ", (int)(t - p->ratesSampling));
      for (tot = 0; tot < p->CFT * p->subPopNumber; tot += p->subPopNumber)
        for (sp = 0; sp < p->subPopNumber; sp++)
          fprintf(n->fp_rates, "%.7g ", ((float)(n->firingsPerPop[tot + sp]) * 1000.0) /
                                            ((float)(n->sp[sp].count) * (float)(p->ratesSampling)));
      fprintf(n->fp_rates, " \n");
      for (k = 0; k < p->CFT * p->subPopNumber; k++) n->firingsPerPop[k] = 0;
    }
  }
  if (x999) {
    if (plastic) commit_plasticity(n);
    if (n->fp_spikes && stat_window(p, sec)) { fclose(n->fp_spikes); n->fp_spikes = NULL; }
    if (n->fp_rates && stat_window(p, sec)) { fclose(n->fp_rates); n->fp_rates = NULL; }
    if (n->fp_probeA) { fclose(n->fp_probeA); n->fp_probeA = NULL; }
    if (n->fp_probeB) { fclose(n->fp_probeB); n->fp_probeB = NULL; }
  }
  if (n->fp_spikes) fflush(n->fp_spikes);
  if (n->fp_rates) fflush(n->fp_rates);
  n->ms = t + 1;
  return 0;
}

uint32_t ora_last_spikes(const ora_net *n, uint32_t *ids, double *te) {
  uint32_t slot = (n->ms + RING - 1) % RING, c = n->ms ? n->ring_cnt[slot] : 0;
  if (ids) memcpy(ids, n->ring_ids[slot], c * 4);
  if (te) memcpy(te, n->ring_te[slot], c * 8);
  return c;
}
void ora_get_state(const ora_net *n, double *v, double *c, double *Tr, double *Te, uint8_t *fired) {
  uint64_t N = n->p.N;
  if (v) memcpy(v, n->v, N * 8);
  if (c) memcpy(c, n->c, N * 8);
  if (Tr) memcpy(Tr, n->Tr, N * 8);
  if (Te) memcpy(Te, n->Te, N * 8);
  if (fired) memcpy(fired, n->didItFire, N);
}
void ora_record_external(ora_net *n, int enable) { n->rec_ext = enable; }
uint32_t ora_last_external(const ora_net *n, uint32_t *counts, double *times, double *values) {
  if (!n->rec_ext) return 0;
  if (counts && n->ext_cnt) memcpy(counts, n->ext_cnt, (uint64_t)n->p.N * 4);
  if (times) memcpy(times, n->ext_t, (uint64_t)n->ext_n * 8);
  if (values) memcpy(values, n->ext_v, (uint64_t)n->ext_n * 8);
  return n->ext_n;
}
double ora_libm_exp(double x) { return exp(x); }
float ora_libm_logf(float x) { return logf(x); }

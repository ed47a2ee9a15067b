"""The product's host side (generator, delivery layout, replayed drive,
parameter interface) and the C ABI surface -- no device needed."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import OracleNet, golden, sha256_file

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def c0_pair(pkg, oracle):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    host = pkg.HostConnectome(cfg)
    ora = OracleNet(oracle, cfg)
    yield host, ora
    host.close()
    ora.close()


def test_library_exports_every_declared_symbol(pkg):
    declared = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            declared |= set(re.findall(r"\b((?:dpsnn|localNetProcess)_\w+)\s*\(", text))
    lib = ctypes.CDLL(pkg.LIB_PATH)
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, missing
    # the Python binding covers the whole dpsnn_* surface
    assert {d for d in declared if d.startswith("dpsnn_")} == set(pkg.exported_symbols())
    assert {"localNetProcess_init", "localNetProcess_fire", "localNetProcess_finish"} <= declared


def test_generator_dump_matches_reference(c0_pair, tmp_path):
    host, _ = c0_pair
    g = golden("c0_synapses.json")
    path = str(tmp_path / "dump.dat")
    host.write_synapse_dump(path)
    assert sha256_file(path) == g["sha256"]


def test_backward_list_matches_oracle(c0_pair):
    host, ora = c0_pair
    a, b = host.synapses(), ora.backward()
    for k in ("pre", "post", "delay", "weight"):
        assert np.array_equal(a[k], b[k]), k


def _reference_delivery_order(b, D):
    """Per target, the (delay, pre, synapse) items in the order the
    reference's delivery loop (DPSNN_localNet_sim.c:858-979) appends them
    when every source fires: delays D-1..0, sources ascending, each run from
    offsetByDelay continuing while the source id matches."""
    pre, post, delay = b["pre"], b["post"], b["delay"]
    S = len(pre)
    items = []
    bstart = np.searchsorted(delay, np.arange(81))
    for d in range(D - 1, -1, -1):
        s, end = bstart[d], bstart[d + 1]
        while s < end:
            src = pre[s]
            e = s
            while e < S and pre[e] == src:
                items.append((post[e], d, src, e))
                e += 1
            s = e if e < end else end
    items.sort(key=lambda x: x[0])  # stable: keeps the append order per target
    return items


def test_inlists_follow_reference_delivery_order(c0_pair, pkg):
    host, ora = c0_pair
    b = ora.backward()
    D = pkg.get_parameters(pkg.config_dir("c0"))["D"]
    want = _reference_delivery_order(b, D)
    il = host.inlists()
    off = il["off"]
    assert int(off[-1]) == len(want)
    tgt = np.repeat(np.arange(host.n), np.diff(off).astype(np.int64))
    got_t = tgt
    want = np.array(want, dtype=np.int64)
    assert np.array_equal(got_t, want[:, 0])
    assert np.array_equal(il["delay"].astype(np.int64), want[:, 1])
    assert np.array_equal(il["pre"].astype(np.int64), want[:, 2])
    assert np.array_equal(il["syn"].astype(np.int64), want[:, 3])
    assert np.array_equal(il["weight"], b["weight"][il["syn"]])


def test_external_parameters_match_oracle(c0_pair):
    host, ora = c0_pair
    a_inv, a_max = host.external_params()
    b_inv, b_max = ora.external_params()
    assert np.array_equal(a_inv, b_inv) and np.array_equal(a_max, b_max)
    # 1000 / (NuExt * CExt) = 1000 / (3 * 800) ms; cap = ceil(20 / invNuExt)
    assert np.all(a_inv == 1000.0 / float(np.float32(3.0) * np.float32(800.0))) and np.all(a_max == 48)


def test_replayed_external_stream_matches_oracle(pkg, oracle):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    host = pkg.HostConnectome(cfg)
    ora = OracleNet(oracle, cfg)
    ora.record_external(True)
    for _ in range(8):
        off, t, v = host.external_step()
        ora.fire()
        cnt, ot, ov = ora.last_external()
        assert np.array_equal(np.diff(off.astype(np.int64)), cnt.astype(np.int64))
        assert np.array_equal(t, ot) and np.array_equal(v, ov)
    host.close()
    ora.close()


@pytest.mark.parametrize("var", ["env_globalSeed", "env_D", "env_CFX", "env_minSynWeight_f",
                                 "env_totalSimTime_ms", "env_spikes_writeDisable"])
def test_missing_env_var_is_a_config_error(pkg, var):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    old = os.environ.pop(var)
    try:
        with pytest.raises(pkg.DpsnnError) as e:
            pkg.get_parameters(cfg)
        assert e.value.code == -1 and var in str(e.value)
    finally:
        os.environ[var] = old


def test_bad_ini_files_are_config_errors(pkg, tmp_path):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    for name in ("column.txt", "connectivity.txt", "bathEfficacy.txt"):
        d = tmp_path / name.split(".")[0]
        d.mkdir()
        for f in ("column.txt", "connectivity.txt", "bathEfficacy.txt"):
            if f != name:
                (d / f).write_text(open(os.path.join(cfg, f)).read())
        with pytest.raises(pkg.DpsnnError):
            pkg.HostConnectome(str(d))
    # a stencil row of the wrong length
    d = tmp_path / "badconn"
    d.mkdir()
    for f in ("column.txt", "bathEfficacy.txt"):
        (d / f).write_text(open(os.path.join(cfg, f)).read())
    (d / "connectivity.txt").write_text("0 0 0.6 0.1 1 10 0.5 0.2\n")
    with pytest.raises(pkg.DpsnnError) as e:
        pkg.HostConnectome(str(d))
    assert "stencil" in str(e.value)


def test_world_size_above_one_is_rejected(pkg):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_globH=2)
    try:
        with pytest.raises(pkg.DpsnnError) as e:
            pkg.get_parameters(cfg)
        assert "env_globH" in str(e.value)
    finally:
        pkg.apply_env(cfg)


def test_parameters_mirror_get_parameters(pkg):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    p = pkg.get_parameters(cfg)
    assert p["globN"] == 5000 and p["neuronsPerCM"] == 1250 and p["subPopNumber"] == 3
    assert p["tabNeuron"] == 1250 and p["D"] == 20 and p["delayMin"] == 1 and p["delayMax"] == 10
    # factors of DPSNN_getParameters.c:378-379, computed in float
    assert p["factorFloat_2_weightType"] == np.float32(-32767) / np.float32(-6.0)
    assert p["factorWeightType_2_Float"] == np.float32(-6.0) / np.float32(-32767)


def _small_grid_config(tmp_path, pkg, grid=4):
    src = pkg.config_dir("c1")  # interior-normalised stencil, as for large grids
    d = tmp_path / f"grid{grid}"
    d.mkdir()
    (d / "column.txt").write_text("0 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n1 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                                  "2 25 0.5 0.0 3 10 20 15 2 0 500 0\n")
    for f in ("connectivity.txt", "bathEfficacy.txt"):
        (d / f).write_text(open(os.path.join(src, f)).read())
    pkg.apply_env(pkg.config_dir("c0"), env_CFX=grid, env_CFY=grid)
    return str(d)


def test_column_parallel_generator_matches_oracle_on_many_columns(pkg, oracle, tmp_path, monkeypatch):
    """36 columns: the generator resolves NormDev's cached value across column
    boundaries exactly like the sequential reference, on one thread (each
    column from its true entry state), on three (a small pool generating both
    entry variants ahead) and on all the host's threads; the in-lists, built
    from pieces of the delivery order on as many threads, keep that order."""
    cfg = _small_grid_config(tmp_path, pkg, grid=6)
    ora = OracleNet(oracle, cfg)
    b = ora.backward()
    want = np.array(_reference_delivery_order(b, pkg.get_parameters(cfg)["D"]), dtype=np.int64)
    # the stream after generation continues identically (external drive)
    ora.record_external(True)
    ext = []
    for _ in range(3):
        ora.fire()
        cnt, ot, ov = ora.last_external()
        ext.append((ot.copy(), ov.copy()))
    ora.close()
    for threads in ("1", "3", None):
        if threads is None:
            monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
        else:
            monkeypatch.setenv("OMP_NUM_THREADS", threads)
        host = pkg.HostConnectome(cfg)
        a = host.synapses()
        for k in ("pre", "post", "delay", "weight"):
            assert np.array_equal(a[k], b[k]), (threads, k)
        il = host.inlists()
        tgt = np.repeat(np.arange(host.n), np.diff(il["off"]).astype(np.int64))
        assert np.array_equal(tgt, want[:, 0]), threads
        assert np.array_equal(il["delay"].astype(np.int64), want[:, 1]), threads
        assert np.array_equal(il["pre"].astype(np.int64), want[:, 2]), threads
        assert np.array_equal(il["syn"].astype(np.int64), want[:, 3]), threads
        for ot, ov in ext:
            off, t, v = host.external_step()
            assert np.array_equal(t, ot) and np.array_equal(v, ov), threads
        host.close()
    pkg.apply_env(pkg.config_dir("c0"))


def test_generator_matches_oracle_on_a_sparse_stencil(pkg, oracle, tmp_path, monkeypatch):
    """36 columns whose stencil mixes a dense centre (q = 0.3) with sparse
    sides (0.01) and a diagonal barely above the reference's 1e-5 floor
    (0.0003): most gap draws land past the sub-population's end, so the
    generator's search of the accumulated gap distribution runs over its
    whole range, against the oracle's linear scan."""
    cfg = _small_grid_config(tmp_path, pkg, grid=6)
    rows = []
    for line in open(os.path.join(cfg, "connectivity.txt")):
        v = line.split()
        if len(v) >= 10:
            rows.append(" ".join(v[:6] + ["0.3", "0.01", "0.01", "0.0003"]))
    (tmp_path / "grid6" / "connectivity.txt").write_text("\n".join(rows) + "\n")
    ora = OracleNet(oracle, cfg)
    b = ora.backward()
    ora.record_external(True)
    ora.fire()
    _, ot, ov = ora.last_external()
    ora.close()
    for threads in ("1", None):
        if threads is None:
            monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
        else:
            monkeypatch.setenv("OMP_NUM_THREADS", threads)
        host = pkg.HostConnectome(cfg)
        a = host.synapses()
        assert len(a["pre"]) > 0
        for k in ("pre", "post", "delay", "weight"):
            assert np.array_equal(a[k], b[k]), (threads, k)
        _, t, v = host.external_step()
        assert np.array_equal(t, ot) and np.array_equal(v, ov), threads
        host.close()
    pkg.apply_env(pkg.config_dir("c0"))


def _check_window(pkg, host, units_per_col):
    """The stencil-window layout streams, per (unit, pass, wave), each owned
    target's in-list items of the pass's run delay in in-list order, padded
    to whole chunks; concatenated over the passes they are the in-list."""
    L = host.window(units_per_col)
    inl = host.inlists()
    off, pre, dly, w = inl["off"], inl["pre"], inl["delay"], inl["weight"]
    slots, npc, passes, unit, upc = L["slots"], L["npc"], L["passes"], L["unit"], L["units_per_col"]
    item, desc, chunk, cols = L["item"], L["desc"], L["chunk"], L["cols"]
    ncols = host.n // npc
    assert L["units"] == ncols * upc and unit * upc >= npc and unit <= 1024
    assert passes == int(dly.max()) + 1
    pad = np.uint32((slots * npc) << 16)
    # every column's sources lie in its stencil slots
    for col in range(ncols):
        srcs = np.unique(pre[off[col * npc]:off[(col + 1) * npc]] // npc)
        assert set(srcs) <= set(cols[col * slots:(col + 1) * slots][cols[col * slots:(col + 1) * slots] >= 0])
    seen = np.zeros(host.n, np.int64)  # items streamed per target
    for u in range(L["units"]):
        col, part = divmod(u, upc)
        j0 = col * npc + part * unit
        m = min(unit, (col + 1) * npc - j0)
        stencil = cols[col * slots:(col + 1) * slots]
        for p in range(passes):
            d = passes - 1 - p
            for wv in range(16):
                s = (u * passes + p) * 16 + wv
                ch = int(chunk[s])
                for q in range(wv, m, 16):
                    j = j0 + q
                    b, e = int(off[j]), int(off[j + 1])
                    seg = np.arange(b, e)[dly[b:e] == d]
                    nch = (len(seg) + 7) // 8
                    assert ((desc[ch:ch + nch] & 0x7FFF) == q).all()
                    # the target's first chunk in the stream is flagged, the others not
                    assert nch == 0 or (desc[ch] & 0x8000 and not (desc[ch + 1:ch + nch] & 0x8000).any())
                    got = item[ch * 8:ch * 8 + len(seg)]
                    widx = got >> 16
                    src = stencil[widx // npc] * npc + widx % npc
                    assert np.array_equal(src, pre[seg]), (u, p, q)
                    assert np.array_equal((got & 0xFFFF).astype(np.uint16).view(np.int16), w[seg])
                    assert (item[ch * 8 + len(seg):(ch + nch) * 8] == pad).all()
                    seen[j] += len(seg)
                    ch += nch
                assert ch == int(chunk[s + 1])
    assert np.array_equal(seen, np.diff(off).astype(np.int64))


def test_window_layout_streams_the_inlists(c0_pair, pkg):
    host, _ = c0_pair
    _check_window(pkg, host, 1)   # units of 625 targets (at most 1024 per unit)
    _check_window(pkg, host, 5)   # units of 250


def test_window_layout_on_grid_edges(pkg, tmp_path):
    """A 3x3 grid of small columns: corner, edge and centre columns have
    stencils of 4, 6 and 9 columns."""
    src = pkg.config_dir("c1")
    d = tmp_path / "g3"
    d.mkdir()
    (d / "column.txt").write_text("0 40 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                                  "1 40 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                                  "2 20 0.5 0.0 3 10 20 15 2 0 500 0\n")
    for f in ("connectivity.txt", "bathEfficacy.txt"):
        (d / f).write_text(open(os.path.join(src, f)).read())
    pkg.apply_env(src, env_CFX=3, env_CFY=3)
    host = pkg.HostConnectome(str(d))
    try:
        L = host.window(2)
        assert L["slots"] == 9
        sizes = sorted(int((L["cols"][c * 9:(c + 1) * 9] >= 0).sum()) for c in range(9))
        assert sizes == [4, 4, 4, 4, 6, 6, 6, 6, 9]
        _check_window(pkg, host, 2)
    finally:
        host.close()


def test_perseo_ini_files(pkg, tmp_path):
    """writeIniFiles (DPSNN_writeIniFiles.c:31-158) on the 2x2-column network: every
    PerseoConFile probability equals the matrix the reference prints at
    init (its stdout, "%f"; rows y, columns x), every column pair within the
    stencil appears once per sub-population pair, sources in column order
    (x outer, y inner); PerseoModFile carries column.txt's values and the
    bath efficacy.  Format strings follow the source; no reference run
    writes these files (its call is compiled out), so the text itself is
    parity unpinned."""
    g = golden("small_golden.json")["djext"]["prob_matrices"]
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    pkg.write_ini_files(cfg, str(tmp_path))
    par = pkg.get_parameters(cfg)
    CFX, CFY, nsub = par["globCFX"], par["globCFY"], par["subPopNumber"]
    half = len(g["0,0"]) // 2
    conn = {}
    for line in open(os.path.join(cfg, "connectivity.txt")):
        f = line.split()
        conn[(int(f[0]), int(f[1]))] = [float(np.float32(x)) for x in f[2:6]]
    lines = open(tmp_path / "PerseoConFile_h0.txt").read().splitlines()
    seen, src_order = [], []
    for ln in lines:
        f = ln.split()
        tp, sp_, prob, dmin, dmax, kind, J, DJ = int(f[0]), int(f[1]), float(f[2]), float(f[3]), float(f[4]), f[5], \
            float(f[6]), float(f[7])
        (tc, ts), (sc, ss) = divmod(tp, nsub), divmod(sp_, nsub)
        dx, dy = tc % CFX - sc % CFX, tc // CFX - sc // CFX
        ref = float(g[f"{ss},{ts}"][dy + half][dx + half])
        assert abs(prob - ref) <= 6e-7 and kind == "'Fixed'", ln
        Jr, DJr, dminr, dmaxr = conn[(ss, ts)]
        assert (dmin, dmax) == (dminr + 0.5, dmaxr + 0.5) and abs(J / Jr - 1) < 1e-5 and abs(DJ / DJr - 1) < 1e-5, ln
        seen.append((sc, tc, ss, ts))
        if not src_order or src_order[-1] != sc:
            src_order.append(sc)
    want = set()
    for scx in range(CFX):
        for scy in range(CFY):
            for tcx in range(CFX):
                for tcy in range(CFY):
                    dx, dy = tcx - scx, tcy - scy
                    if abs(dx) > half or abs(dy) > half:
                        continue
                    for ss in range(nsub):
                        for ts in range(nsub):
                            if float(g[f"{ss},{ts}"][dy + half][dx + half]) > 0:
                                want.add((scx + scy * CFX, tcx + tcy * CFX, ss, ts))
    assert len(seen) == len(set(seen)) and set(seen) == want
    assert src_order == [cx + cy * CFX for cx in range(CFX) for cy in range(CFY)]
    col = [ln.split() for ln in open(os.path.join(cfg, "column.txt")) if ln.strip()]
    bath = float(open(os.path.join(cfg, "bathEfficacy.txt")).read().split()[0])
    mods = open(tmp_path / "PerseoModFile_h0.txt").read().splitlines()
    assert len(mods) == CFX * CFY * nsub
    for k, ln in enumerate(mods):
        f, c = ln.split(), col[k % nsub]
        assert ln.endswith(" 0 ") and int(f[0]) == int(c[1]) and float(f[3]) == bath
        # JExt DJExt | NuExt Tau Theta H Tarp AlphaC TauC gC
        for got, want_ in zip(f[1:3] + f[4:12], c[2:4] + c[4:12]):
            assert abs(float(got) - float(want_)) <= 1e-6 * max(1.0, abs(float(want_))), (ln, got, want_)


def test_prepared_network_reports_host_and_device_bytes(pkg):
    """A network built on the host only (dpsnn_prepare, no device call)
    reports the host bytes its synapse and delivery lists hold -- what the
    server's backlog bound counts -- and the device bytes its upload would
    take; it can be destroyed without ever being uploaded."""
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    lib = pkg.library()
    opts = pkg.Options()
    opts.drive = pkg.DRIVE_REPLAY
    h = ctypes.c_void_p()
    rc = lib.dpsnn_prepare(cfg.encode(), ctypes.byref(opts), ctypes.byref(h))
    assert rc == 0, lib.dpsnn_last_error().decode()
    try:
        S = 5621357  # the reference's synapse count for this network
        host = lib.dpsnn_host_bytes(h)
        # backward list (12 B per synapse) plus the in-lists (10 B per item
        # and a 2-byte local index) before upload
        assert 22 * S <= host <= 40 * S, host
        assert lib.dpsnn_device_bytes_needed(h) > 6 * S
    finally:
        lib.dpsnn_destroy(h)

"""GPU parity of the remaining input paths against the FP64 oracle, through
the C ABI:

* every target reradiation function (rerad_funct 0-4, radar_kernels.cu:
  161-195) of type-0 targets (genRefrPhasor, radar_kernels.cu:351-398) on the
  specular linear-convolution path, the specular direct-sum path
  (refrRadarSignal, radar_kernels.cu:201-299) and the plain linear path;
* per-shot tables read from files through Simulation.from_files (sim.cu:
  624-651): a source path with its own normals, a gain file and an rx-window
  file whose values change from shot to shot, plus an attenuation prism
  file (file_io.cu:429-624); once on the plain method with typed targets and
  once on the specular circular convolution, whose chirp table depends on
  each shot's rx window start (genChirp, radar_kernels.cu:454-466);
* the drop-in CLI `python -m europaraypaths_amd.sim` end to end: s%06d.txt
  per lit shot, none for a dark one, read back with compile_rdrgrm.

Tolerance as in test_gpu_parity.py (tests/helpers.py::compare_traces).
"""
import json
import os

import numpy as np
import pytest

from europaraypaths_amd import Simulation, formats, surfaces, workloads
from europaraypaths_amd.output_handling import compile_rdrgrm
from europaraypaths_amd.params import make_shots, parse_parameters
from oracle import oracle
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu


def _scene(nx, ny, fs, nr, h, rst, depth, **over):
    ox, oy = -fs * nx / 2, -fs * ny / 2
    z = surfaces.gaussian_heights(nx, ny, fs, 0.3, 8.0 * fs, seed=3)
    soa = surfaces.terrain_soa(nx, ny, fs, ox, oy, z)
    p = workloads.make_params("REASON_VHF", "planetary_ice", nr=nr, smpl=48e6, rst=rst,
                              aperture=30.0, fs=fs, ns=2, sx0=-0.3 * fs, sdx=0.6 * fs,
                              sy=0.2 * fs, sz=h, **over)
    par = parse_parameters(p)
    xyzn = np.zeros((6, 3), np.float32)
    xyzn[0] = [0.0, 1.3 * fs, -1.1 * fs]
    xyzn[1] = [0.4 * fs, -0.8 * fs, 0.9 * fs]
    xyzn[2] = -depth
    xyzn[5] = 1.0
    return par, soa, (xyzn, np.zeros(3, np.uint8)), make_shots(par)


# (method, mesh facets per side, facet size, nr, altitude, rx start, depth,
#  parameter overrides).  Facets of 2.5-5 m put several facets inside the
#  0.5 deg and 1 deg cones of the boxcar functions at these depths.
_METHODS = {
    "specular_linear": (64, 5.0, 1024, 3000.0, 2600.0, 600.0, {"specular": True}),
    "specular_noconv": (64, 2.5, 256, 3000.0, 3400.0, 300.0,
                        {"specular": True, "convolution": False}),
    "plain_linear": (64, 5.0, 1024, 3000.0, 2600.0, 600.0, {}),
}


@pytest.mark.parametrize("method", sorted(_METHODS))
@pytest.mark.parametrize("rerad", [0, 1, 2, 3, 4])
def test_type0_reradiation_functions(gpu_ctx, method, rerad):
    n, fs, nr, h, rst, depth, over = _METHODS[method]
    print(f"type-0 targets {method} rerad_funct={rerad}", flush=True)
    par, soa, tg, shots = _scene(n, n, fs, nr, h, rst, depth, rerad_funct=rerad, **over)
    e = par.to_erp()
    assert e.rerad_funct == rerad
    o, ov = oracle.run(e, soa, tg, shots)
    # the boxcars must keep some facets and drop others, or they test nothing
    assert np.abs(o).max() > 0
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        g, gv = gpu_ctx.run(e, m, t, shots)
    finally:
        m.release()
        t.release()
    np.testing.assert_array_equal(gv, ov)
    w = compare_traces(g, o)
    print(f"worst errors {w}", flush=True)


def test_reradiation_functions_differ(gpu_ctx):
    """The five functions give five different traces on one scene (so the
    parametrised case above cannot pass on a kernel that ignores the
    selector), except 3, which the reference's degree/radian slip makes
    equal to 0."""
    n, fs, nr, h, rst, depth, over = _METHODS["specular_linear"]
    outs = {}
    for rerad in range(5):
        par, soa, tg, shots = _scene(n, n, fs, nr, h, rst, depth, rerad_funct=rerad, **over)
        m = gpu_ctx.upload_mesh(soa)
        t = gpu_ctx.upload_targets(*tg)
        try:
            outs[rerad], _ = gpu_ctx.run(par.to_erp(), m, t, shots)
        finally:
            m.release()
            t.release()
    np.testing.assert_allclose(outs[3], outs[0], rtol=0, atol=1e-6 * np.abs(outs[0]).max())
    for a, b in ((0, 1), (1, 2), (0, 4), (2, 4)):
        d = np.abs(outs[a] - outs[b]).max() / np.abs(outs[a]).max()
        assert d > 1e-2, (a, b, d)


def _write_inputs(tmp, specular=False, dark_shot=None):
    """A small scene written in the reference's file formats with per-shot
    source positions and normals, gains and rx windows that change from
    shot to shot, and one attenuation prism."""
    nx = ny = 40
    fs = 10.0
    ox, oy = -fs * nx / 2, -fs * ny / 2
    z = surfaces.gaussian_heights(nx, ny, fs, 0.4, 60.0, seed=9)
    soa = surfaces.terrain_soa(nx, ny, fs, ox, oy, z)
    formats.export_facets(*soa, os.path.join(tmp, "facets.fct"))
    xs = np.array([-60.0, 0.0, 55.0, 20.0])
    ys = np.array([10.0, -30.0, 0.0, 35.0])
    ttype = 0 if specular else 1
    with open(os.path.join(tmp, "targets.txt"), "w") as f:
        f.write("\n".join(f"{x:.6f},{y:.6f},-250.000000,0.000000,0.000000,1.000000,{ttype}"
                          for x, y in zip(xs, ys)))
    ns = 5
    src = []
    for s in range(ns):
        tilt = 0.03 * (s - 2)
        nrm = np.array([tilt, -0.5 * tilt, 1.0])
        nrm /= np.linalg.norm(nrm)
        if dark_shot is not None and s == dark_shot:
            nrm = np.array([0.0, 0.0, -1.0])  # antenna pointing up: nothing lit
        src.append((-40.0 + 20.0 * s, 5.0 * s, 3000.0 + 35.0 * s, *nrm))
    with open(os.path.join(tmp, "sources.txt"), "w") as f:
        f.write("\n".join(",".join(f"{v:.6f}" for v in r) for r in src))
    with open(os.path.join(tmp, "gains.txt"), "w") as f:
        f.write("\n".join(f"{3.0 + 0.7 * s:.3f},{5.0 - 0.9 * s:.3f}" for s in range(ns)))
    with open(os.path.join(tmp, "rxwin.txt"), "w") as f:
        for s in range(ns):
            f.write(f"{2700.0 + 35.0 * s - 11.0 * (s % 2)}\n")
    with open(os.path.join(tmp, "prisms.txt"), "w") as f:
        f.write("0.00002,-100,-100,-200,20,100,-50\n")
    rec = {"rx_window_m": workloads.rx_window_for(512 if specular else 1024, 48e6),
           "rx_sample_rate": 48e6, "rx_window_position_file": "rxwin.txt"}
    src_par = {"ns": ns, "aperture": 25.0, "source_path_file": "sources.txt"}
    over = {"gain_pattern_file": "gains.txt", "attenuation_geometry_file": "prisms.txt",
            "sig_2": 1e-5}
    if specular:
        over.update({"specular": True, "convolution_linear": False})
    from europaraypaths_amd.params import gen_params
    p = gen_params("REASON_VHF", "planetary_ice", {"fs": fs}, rec, src_par, par=over)
    with open(os.path.join(tmp, "params.json"), "w") as f:
        json.dump(p, f, indent=4)
    return [os.path.join(tmp, n) for n in ("params.json", "facets.fct", "targets.txt")], p


@pytest.mark.parametrize("specular", [False, True])
def test_per_shot_tables_from_files(gpu_ctx, tmp_path, specular):
    files, _ = _write_inputs(str(tmp_path), specular=specular)
    sim = Simulation.from_files(*files)
    sh = sim.shots
    # the tables really change from shot to shot
    for k in ("rst", "grefl_lin", "grefr_lin", "sz", "snx"):
        assert np.unique(sh[k]).size == sh.shape[0], k
    assert sim.prisms.shape[0] == 1
    print(f"per-shot tables specular={specular}", flush=True)
    rdr, gv = sim.run(ctx=gpu_ctx)
    o, ov = oracle.run(sim.par.to_erp(), sim.mesh, sim.targets, sh, prisms=sim.prisms)
    np.testing.assert_array_equal(gv, ov)
    assert (gv > 0).all()
    w = compare_traces(rdr.T, o)
    print(f"worst errors {w}", flush=True)


def test_cli_end_to_end(gpu_ctx, tmp_path):
    """python -m europaraypaths_amd.sim on files in the reference's formats:
    one s%06d.txt per lit shot, none for the dark shot, the text traces read
    back with compile_rdrgrm equal to the oracle's radargram without that
    column (the reference's column shift), rdrgrm.npy the full radargram."""
    from europaraypaths_amd import sim as sim_cli
    files, p = _write_inputs(str(tmp_path), dark_shot=3)
    out = tmp_path / "out"
    out.mkdir()
    (out / "keep.dat").write_text("not a trace")
    (out / "s999999.txt").write_text("stale trace from an earlier run")
    rc = sim_cli.main([*files, str(out), "--npy"])
    assert rc == 0
    names = sorted(os.listdir(out))
    assert names == ["keep.dat", "rdrgrm.npy", "s000000.txt", "s000001.txt", "s000002.txt",
                     "s000004.txt"], names
    sim = Simulation.from_files(*files)
    o, ov = oracle.run(sim.par.to_erp(), sim.mesh, sim.targets, sim.shots, prisms=sim.prisms)
    assert ov[3] == 0 and (np.delete(ov, 3) > 0).all()
    rdr_txt = compile_rdrgrm(str(out), p)
    assert rdr_txt.shape == (sim.par.nr, 4)
    compare_traces(rdr_txt.T, np.delete(o, 3, axis=0))
    rdr_npy = np.load(out / "rdrgrm.npy")
    assert rdr_npy.shape == (sim.par.nr, 5)
    assert np.all(rdr_npy[:, 3] == 0)
    compare_traces(rdr_npy.T, o)
    # "%e" keeps 7 significant digits of each float32 component
    np.testing.assert_allclose(rdr_txt, np.delete(rdr_npy, 3, axis=1), rtol=1e-5,
                               atol=1e-6 * np.abs(rdr_npy).max())

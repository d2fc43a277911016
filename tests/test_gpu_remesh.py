"""Stage 1 (remeshing) on the GPU against the oracle's restatement: the
signed distance lattice and the extracted mesh bit for bit, and PaMO.run
with stage 1 on."""
import numpy as np
import pytest
import torch

import oracle
from pamo_amd import PaMO, remesh, synth
from pamo_amd import verify as pv

pytestmark = pytest.mark.gpu


def _cases():
    return [("icosphere4", synth.icosphere(4)), ("bumpy", synth.displaced_icosphere(4, 3)),
            ("torus", synth.noisy_torus(40, 20, 5)), ("sdf_union", synth.sdf_union(12, 48, 9))]


@pytest.mark.parametrize("R", [32, 77, 128])
def test_sdf_and_extraction_exact(cuda, R):
    """get_sdf equals the oracle's lattice and the extraction of the shifted
    lattice equals the oracle's mesh, bit for bit; the shell is closed."""
    band = 3.0 / R
    margin = 2 * band + 1
    for name, (v, f) in _cases():
        tris, tmin, tmax, mean = remesh.normalize(v, f, band, margin)
        print("remesh", name, R, f.shape, flush=True)
        vol = remesh.get_sdf(torch.from_numpy(tris).to(cuda), R, band).cpu().numpy()
        ref = oracle.mesh_to_sdf(tris, R, band)
        assert np.array_equal(vol.view(np.uint32), ref.view(np.uint32)), name
        d = vol - np.float32(0.9 / R)
        gv, gf = remesh.VolumeToMesh()(torch.from_numpy(d).to(cuda))
        ov, of = oracle.volume_to_mesh(d)
        assert np.array_equal(gv.cpu().numpy().view(np.uint32), ov.view(np.uint32)), name
        assert np.array_equal(gf.cpu().numpy(), of), name
        r = pv.verify(ov, of)
        assert r["watertight"] and r["manifold"] and r["misoriented_edges"] == 0, (name, r)


def test_large_lattice_closed(cuda):
    """R = 256, the reference's default lattice, on a 130k-face union:
    closed, manifold, oriented."""
    v, f = synth.sdf_union(60, 160, 4)
    R = 256
    band = 3.0 / R
    tris, tmin, tmax, mean = remesh.normalize(v, f, band, 2 * band + 1)
    gv, gf = remesh.remesh(torch.from_numpy(tris).to(cuda), tmin, tmax, R, band, 2 * band + 1)
    print("R 256", f.shape, gv.shape, gf.shape, flush=True)
    r = pv.verify(gv.cpu().numpy(), gf.cpu().numpy())
    assert r["watertight"] and r["manifold"] and r["misoriented_edges"] == 0, r


def test_pamo_run_with_stage1(cuda):
    """PaMO(use_stage1=True).run: the remeshed shell (a closed manifold 0.9
    lattice steps outside the input) simplified by stage 2 to its target,
    identical to stage 2 alone on the same remeshed mesh."""
    from pamo_amd import simplify_stage2, _scale
    v, f = synth.displaced_icosphere(5, 2)
    pm = PaMO((v, f), use_stage1=True, use_stage3=False, verbose=False)
    # (min_verts=0: the reference's default, 1e10, makes the target the
    # whole mesh, and stage 2 stops after its first round)
    ov, of = pm.run(torch.from_numpy(v).to(cuda), torch.from_numpy(f).to(cuda), 0.05, min_verts=0)
    assert pm.R == 256 and pm.last_stage1_seconds is not None
    r = pv.verify(ov.astype(np.float32), of)
    assert r["watertight"] and r["manifold"], r
    assert len(of) <= max(int(0.05 * len(f)), 10) or pm.last_stats["exit_reason"] == 3
    # the same remeshed mesh through stage 2 alone
    band, margin = 3 / 256, 2 * (3 / 256) + 1
    tris, tmin, tmax, mean = remesh.normalize(v, f, band, margin)
    rv, rf = remesh.remesh(torch.from_numpy(tris).to(cuda), tmin, tmax, 256, band, margin)
    sv, sf, st = simplify_stage2(rv, rf, _scale(rv), int(0.05 * len(f)))
    assert np.array_equal(sf.cpu().numpy(), of)
    assert np.array_equal((sv.cpu().numpy() + mean).view(np.uint32), ov.astype(np.float32).view(np.uint32))

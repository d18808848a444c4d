"""The `run` stage end to end on a synthetic work directory (the reference's
preprocess outputs: affine_prealign.json, ref_preprocessed.npy,
def_preprocessed.npy, def_ref_aligned.npy; run_dvc.py:65-192): the solve
recovers the known displacement, and the voxel- and element-scale outputs
equal the oracle's evaluation of the same stages on the saved U, dU and
residual."""
import json
import os

import numpy as np
import pytest

from oracle import cells_oracle as CO
from oracle import post_oracle as P

pytestmark = pytest.mark.gpu


def _work_dir(tmp_path, n=65, h=8):
    from voxeldvc_amd import synth
    f = synth.speckle_numpy((n, n, n), 2500, 21)
    t = (1.3, -0.6, 0.4)
    g = synth.warp_backward(f, synth.translation_strain_field(t, 0.004, n / 2), 0.0, 22)
    f8 = np.clip(np.round(f), 0, 255).astype(np.uint8)
    g8 = np.clip(np.round(g), 0, 255).astype(np.uint8)
    wd = str(tmp_path)
    meta = dict(h=h, glt=0.0, affine_A=np.eye(3).tolist(), affine_t=[0.0, 0.0, 0.0],
                affine_c=[n / 2] * 3, crop_lo=[0, 0, 0], crop_hi=[n, n, n], def_lo=[0, 0, 0])
    with open(os.path.join(wd, 'affine_prealign.json'), 'w') as fh:
        json.dump(meta, fh)
    np.save(os.path.join(wd, 'ref_preprocessed.npy'), f8)
    np.save(os.path.join(wd, 'def_preprocessed.npy'), g8)
    np.save(os.path.join(wd, 'def_ref_aligned.npy'), g8)
    return wd, f8, g8, t


def test_run_stage_outputs(gpu, tmp_path):
    from voxeldvc_amd.run_dvc import main
    n, h = 65, 8
    N = (n - 1) // h
    wd, f8, g8, t = _work_dir(tmp_path, n, h)
    print('run stage on', wd, flush=True)
    main([wd])
    U_grid = np.load(os.path.join(wd, 'U_recovered.npy'))
    res = np.load(os.path.join(wd, 'residual.npy'))
    U_local = np.load(os.path.join(wd, 'U_local.npy'))
    assert U_grid.shape == (N + 1, N + 1, N + 1, 3) and res.shape == (n, n, n)
    # recovered displacement at interior nodes: translation + small z strain
    k = np.arange(N + 1) * h
    want_z = t[2] + 0.004 * (k - n / 2)
    inner = slice(1, N)
    assert np.abs(U_grid[inner, inner, inner, 0] - t[0]).max() < 0.1
    assert np.abs(U_grid[inner, inner, inner, 1] - t[1]).max() < 0.1
    assert np.abs(U_grid[inner, inner, inner, 2] - want_z[None, None, inner]).max() < 0.1
    # voxel-scale outputs against the oracle on the saved U / residual
    pushed = np.load(os.path.join(wd, 'ref_forward_pushed.npy'))
    assert pushed.dtype == np.uint8 and pushed.shape == (n, n, n)
    np.testing.assert_array_equal(pushed, P.push_field_forward(f8, U_local, N, N, N, h))
    res_flat = res.reshape(-1, order='F')
    ws, wz = P.compute_correlation_quality_cell(res_flat, f8, h, N, N, N)
    for name, want in (('residual_std_cell', ws), ('zncc_cell', wz)):
        got = np.load(os.path.join(wd, name + '.npy'))
        assert got.shape == (N, N, N)
        np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        np.testing.assert_allclose(got[ok], want[ok], rtol=1e-10, atol=1e-12)
    U = np.zeros(3 * (N + 1) ** 3)
    for c in range(3):
        U[c::3] = U_grid[..., c].ravel(order='F')
    ff = f8.reshape(-1, order='F')
    act = P.compute_active_mask(U, g8, h, N, N, N, mask0_vox=ff > 0.0, glt=0.0)
    want_zncc = P.compute_zncc(ff, ff.astype(np.float64) - res_flat, act)
    log = open(os.path.join(wd, 'run_log.txt')).read()
    line = [ln for ln in log.splitlines() if ln.startswith('ZNCC (active voxels')][0]
    assert float(line.split(':')[-1]) == pytest.approx(want_zncc, abs=2e-6)
    assert want_zncc > 0.95
    assert 'DVC COMPLETE' in log
    # element-scale outputs against the oracle on the same active mask
    uv = np.load(os.path.join(wd, 'unsafe_voxel_mask.npy'))
    assert uv.dtype == np.bool_ and uv.shape == (n, n, n)
    np.testing.assert_array_equal(uv.reshape(-1, order='F'),
                                  CO.unsafe_voxel_mask(act, h, N, N, N))
    ol = np.load(os.path.join(wd, 'overlap_lost_element_mask.npy'))
    assert ol.shape == (N, N, N)
    np.testing.assert_array_equal(ol.reshape(-1, order='F'),
                                  CO.overlap_lost_element_mask(act, ff > 0.0, h, N, N, N))
    ps = np.load(os.path.join(wd, 'principal_strains_cell.npy'))
    want_ps = CO.principal_strains_cell(U_grid, h)
    assert ps.shape == (N, N, N, 3)
    assert np.abs(ps - want_ps).max() <= 1e-13 * np.abs(want_ps).max()
    # the imposed strain: e_zz = 0.004 in the interior
    assert abs(np.median(ps[2:-2, 2:-2, 2:-2, 0]) - 0.004) < 5e-4
    eq = np.load(os.path.join(wd, 'equivalent_strain_cell.npy'))
    np.testing.assert_allclose(eq, CO.equivalent_strain(ps), rtol=2.3e-16, atol=0)
    dU = np.load(os.path.join(wd, 'dU.npy'))
    conv = np.load(os.path.join(wd, 'gn_convergence_cell.npy'))
    np.testing.assert_allclose(conv, CO.gn_convergence_cell(dU, U, N, N, N), rtol=2.3e-16,
                               atol=0)
    od = np.load(os.path.join(wd, 'overlap_lost_element_mask_deformed.npy'))
    np.testing.assert_array_equal(od, CO.remap_mask_nearest(ol, U_grid, h, (0, 0, 0)))
    up = np.load(os.path.join(wd, 'unsafe_voxel_mask_forward_pushed.npy'))
    assert up.dtype == np.bool_
    np.testing.assert_array_equal(
        up, P.push_field_forward(uv, U_local, N, N, N, h, order=0, mode='constant', cval=1))
    # deformed-configuration strains: the device interpolant (the oracle's
    # exhaustive-search restatement of it) and the reference's griddata
    psd = np.load(os.path.join(wd, 'principal_strains_cell_deformed.npy'))
    assert psd.shape == (N, N, N, 3)
    want_psd = CO.remap_field_kuhn_brute(ps, U_grid, h)
    np.testing.assert_allclose(psd, want_psd, rtol=0, atol=1e-12 * np.abs(ps).max())
    grid_psd = CO.remap_field_linear(ps, U_grid, h)
    d2 = max(np.abs(np.diff(ps, 2, axis=a)).max() for a in range(3))
    assert np.abs(psd - grid_psd)[1:-1, 1:-1, 1:-1].max() <= 0.5 * d2 + 1e-12
    eqd = np.load(os.path.join(wd, 'equivalent_strain_cell_deformed.npy'))
    np.testing.assert_allclose(eqd, CO.equivalent_strain(psd), rtol=2.3e-16, atol=0)
    mat = np.load(os.path.join(wd, 'material_cell_mask.npy'))
    np.testing.assert_array_equal(mat, CO.material_cell_mask(f8, h, N, N, N, 0.0))
    mero = np.load(os.path.join(wd, 'material_cell_mask_eroded.npy'))
    np.testing.assert_array_equal(mero, CO.material_cell_mask_eroded(mat))

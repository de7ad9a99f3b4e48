"""The header-only C++ entry points (include/tipl: tipl::reg::linear,
tipl::reg::cdm_common) driven through tests/cpp/reg_cli, built by
__graft_entry__.build() with g++ against libtiplhip.so.

linear: ground truth recovery on the seeded 96x112x96 blob pair (moving =
fixed under a random 12-DOF affine); the reference's own optimiser is
schedule-dependent, so the bar is the transform error against the truth.
cdm: the field equals the C ABI's on the same inputs (same library path).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tests", "cpp", "reg_cli")


def _transform_error(p, truth, shape):
    from tipl_amd import phantoms
    m, sh = phantoms.affine_matrix_np(p, shape, shape)
    mt, sht = phantoms.affine_matrix_np(truth, shape, shape)
    zz, yy, xx = np.meshgrid(*[np.arange(0, n, 2.0) for n in shape], indexing="ij")
    pts = np.stack([xx, yy, zz], 0).reshape(3, -1)
    d = np.linalg.norm((m @ pts + sh[:, None]) - (mt @ pts + sht[:, None]), axis=0)
    return float(d.mean()), float(d.max())


@pytest.fixture(scope="module")
def blob_pair():
    from tipl_amd import phantoms
    fixed, moving, truth = phantoms.blob_pair(0)
    return phantoms.quantize(moving), phantoms.quantize(fixed), truth


@pytest.mark.parametrize("cost", ["mi", "corr"])
def test_linear_recovers_truth(tmp_path, blob_pair, cost):
    assert os.path.exists(CLI), "tests/cpp/reg_cli is built by __graft_entry__.build()"
    frm, to, truth = blob_pair
    a, b, out = tmp_path / "a.u8", tmp_path / "b.u8", tmp_path / "out.txt"
    frm.tofile(a)
    to.tofile(b)
    z, y, x = frm.shape
    r = subprocess.run([CLI, "linear", str(a), str(b), str(x), str(y), str(z), cost, str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    vals = open(out).read().split()
    arg = np.array(vals[:12], np.float32)
    mean_err, max_err = _transform_error(arg, truth, frm.shape)
    print(f"{cost}: transform error mean {mean_err:.4f} max {max_err:.4f} voxel, cost {vals[12]}")
    assert mean_err < 0.5 and max_err < 1.5


def test_cdm_cli_matches_c_abi(tmp_path, ctx):
    from scipy.ndimage import shift
    from conftest import smooth_volume
    It = smooth_volume((40, 44, 36), 21, sigma=2.5)
    Is = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)
    a, b, out = tmp_path / "t.u8", tmp_path / "s.u8", tmp_path / "f.f32"
    It.tofile(a)
    Is.tofile(b)
    z, y, x = It.shape
    r = subprocess.run([CLI, "cdm", str(a), str(b), str(x), str(y), str(z), str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    f_cli = np.fromfile(out, np.float32).reshape(It.shape + (3,))
    rc, f, _ = ctx.cdm(ctx.volume(It), ctx.volume(Is))
    assert rc == 0
    assert np.array_equal(f_cli.view(np.uint32), f.view(np.uint32))


def test_cdm2d_cli_matches_c_abi(tmp_path, ctx):
    """tipl::reg::cdm_common on image<2,uint8> through the C++ mirror equals
    the C ABI's tiplhip_cdm2d bit for bit."""
    from scipy.ndimage import shift
    from test_oracle_cdm2d import image2
    It = image2((90, 110), 22, sigma=2.0)
    Is = shift(It.astype(np.float32), (-1.5, 2.0), order=1).astype(np.uint8)
    a, b, out = tmp_path / "t.u8", tmp_path / "s.u8", tmp_path / "f.f32"
    It.tofile(a)
    Is.tofile(b)
    h, w = It.shape
    r = subprocess.run([CLI, "cdm2d", str(a), str(b), str(w), str(h), str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    f_cli = np.fromfile(out, np.float32).reshape(It.shape + (2,))
    rc, f, _ = ctx.cdm2d(It, Is)
    assert rc == 0 and np.abs(f).max() > 0.1
    assert np.array_equal(f_cli.view(np.uint32), f.view(np.uint32))


@pytest.mark.parametrize("channels,bandwidth", [(1, 1), (3, 8), (1, 3)])
def test_gaussian_cli_matches_oracle(tmp_path, oracle, channels, bandwidth):
    """tipl::filter::gaussian from the C++ mirror equals the oracle bit for bit."""
    rng = np.random.default_rng(channels * 10 + bandwidth)
    shape = (23, 31, 19)
    img = rng.standard_normal(shape + ((3,) if channels == 3 else ())).astype(np.float32) * 50
    a, out = tmp_path / "in.f32", tmp_path / "out.f32"
    img.tofile(a)
    z, y, x = shape
    r = subprocess.run([CLI, "gaussian", str(a), str(x), str(y), str(z), str(channels),
                        str(bandwidth), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(img.shape)
    assert np.array_equal(got.view(np.uint32), oracle.gaussian(img, bandwidth).view(np.uint32))


def test_subject_pre_cli_matches_oracle(tmp_path, oracle):
    from tipl_amd import phantoms
    f, _, _ = phantoms.serving_pair(7, (40, 36, 44), False, np.float32)
    a, out = tmp_path / "in.f32", tmp_path / "out.u8"
    f.tofile(a)
    z, y, x = f.shape
    r = subprocess.run([CLI, "subject", str(a), str(x), str(y), str(z), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(np.fromfile(out, np.uint8).reshape(f.shape), oracle.subject_pre(f))


def test_mm_reg_cli_matches_oracle(tmp_path, oracle):
    """tipl::reg::mm_reg (the C++ mirror) with a given affine: linear_reg ->
    calculate_linear_r and It_match_fov, nonlinear_reg with the reverse cdm
    on its own thread, mappings, calculate_nonlinear_r; every output equals
    the oracle's composition of the same steps bit for bit (fixed-schedule
    cdm, so the whole run is bit-exact)."""
    from scipy.ndimage import shift
    from conftest import smooth_volume
    from mmreg_oracle import oracle_mm_reg
    It = smooth_volume((30, 34, 28), 3, sigma=2.5)
    I = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)[1:, :, 2:]
    arg = np.array([0.5, -0.3, 0.2, 0.02, -0.01, 0.015, 1.0, 1.0, 1.0, 0, 0, 0], np.float32)
    vt, vs = (1.0, 1.0, 1.0), (1.1, 1.0, 0.9)
    param = np.concatenate([arg, vt, vs, [2, 4, 0.5, 2]]).astype(np.float32)
    a, b, p, out = tmp_path / "It.u8", tmp_path / "I.u8", tmp_path / "p.f32", tmp_path / "out"
    It.tofile(a)
    I.tofile(b)
    param.tofile(p)
    tz, ty, tx = It.shape
    sz, sy, sx = I.shape
    r = subprocess.run([CLI, "mmreg", str(a), str(tx), str(ty), str(tz), str(b), str(sx), str(sy),
                        str(sz), str(p), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want = oracle_mm_reg(oracle, It, vt, I, vs, arg, cdm_fixed=(2, 4, 0.5, 2))
    rr = np.fromfile(str(out) + ".r", np.float32)
    assert rr[0] == want["r"][0] and rr[1] == want["r"][1]
    f3 = lambda ext, shp: np.fromfile(str(out) + ext, np.float32).reshape(shp + (3,))
    for ext, key, shp in ((".t2f", "t2f", It.shape), (".f2t", "f2t", It.shape),
                          (".to2from", "to2from", It.shape), (".from2to", "from2to", I.shape)):
        assert np.array_equal(f3(ext, shp).view(np.uint32), want[key].view(np.uint32)), key
    assert np.array_equal(np.fromfile(str(out) + ".JJ", np.uint8).reshape(It.shape), want["JJ"])
    jac = np.fromfile(str(out) + ".jac", np.float32).reshape(It.shape)
    assert np.array_equal(jac.view(np.uint32), oracle.jacobian_determinant_dis(want["t2f"]).view(np.uint32))

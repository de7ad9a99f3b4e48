"""NIfTI reading and writing (host code in libtiplhip; no GPU work).

Parity: every generated file is read by the library and by the reference's
tipl::io::gz_nifti (compiled from the reference headers into oracle/_ref)
and the voxels, dims, voxel sizes, transformation and is_mni must agree bit
for bit; files written by either side are read back by the other.  Without
the reference build the closed-form checks still pin the orientation rules:
an RAS file flips x and y into LPS, a swapped-axis file transposes, scl is
applied after the type conversion, and byte order, compression, NIfTI-2 and
Analyze pairs do not change what is read."""
import numpy as np
import pytest

import nifti_files


def _lib():
    import tipl_amd
    return tipl_amd


def _vol(shape=(5, 6, 7), seed=0, lo=-100, hi=3000):
    r = np.random.default_rng(seed)
    return r.integers(lo, hi, shape).astype(np.float64)


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


CASES = {
    "lps_f32": dict(datatype=16),
    "ras_i16": dict(datatype=4, srow=[1.5, 0, 0, -20, 0, 1.2, 0, -30, 0, 0, 2.0, -10], vs=(1.5, 1.2, 2.0)),
    "lpi_u8": dict(datatype=2, srow=[-1, 0, 0, 5, 0, -1, 0, 6, 0, 0, -1, 7]),
    "swap_xy": dict(datatype=16, srow=[0, -1, 0, 3, -1, 0, 0, 4, 0, 0, 1, 5]),
    "swap_xz": dict(datatype=16, srow=[0, 0, 2, 3, 0, -1, 0, 4, -1.5, 0, 0, 5], vs=(1.5, 1, 2)),
    "swap_yz": dict(datatype=8, srow=[-1, 0, 0, 3, 0, 0, -2, 4, 0, 1.5, 0, 5], vs=(1, 1.5, 2)),
    "oblique": dict(datatype=16, srow=[0.9, 0.3, 0.1, 1, -0.2, 0.95, 0.05, 2, 0.1, -0.1, 1.1, 3]),
    "big_endian_f64": dict(datatype=64, endian=">"),
    "big_endian_i16_ras": dict(datatype=4, endian=">", srow=[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]),
    "scl": dict(datatype=4, scl=(0.5, 10.0)),
    "scl_slope_only": dict(datatype=2, scl=(2.0, 0.0)),
    "uint16": dict(datatype=512),
    "int8": dict(datatype=256),
    "uint32": dict(datatype=768),
    "int64": dict(datatype=1024),
    "qform": dict(datatype=16, sform=0, qform=1, quat=(0.1, -0.2, 0.3), qoff=(4.0, -5.0, 6.0),
                  vs=(1.1, 0.9, 1.3), qfac=-1.0),
    "qform_identity": dict(datatype=16, sform=0, qform=1, quat=(0.0, 0.0, 0.0), qoff=(1, 2, 3)),
    "mni": dict(datatype=16, sform=4, srow=[-1, 0, 0, 90, 0, -1, 0, 126, 0, 0, 1, -72]),
    "nifti2": dict(datatype=16, version=2),
    "nifti2_be_ras": dict(datatype=4, version=2, endian=">", srow=[1, 0, 0, 1, 0, 1, 0, 2, 0, 0, 1, 3]),
    "gz": dict(datatype=4, gz=True),
    "nifti2_gz": dict(datatype=16, version=2, gz=True, srow=[0, -1, 0, 3, -1, 0, 0, 4, 0, 0, 1, 5]),
    "analyze": dict(datatype=4, version=0),
}


def _write_case(tmp_path, name, vol):
    kw = dict(CASES[name])
    ext = ".hdr" if kw.get("version") == 0 else (".nii.gz" if kw.get("gz") else ".nii")
    path = tmp_path / (name + ext)
    lo_hi = {2: (0, 255), 256: (-128, 127), 512: (0, 60000), 768: (0, 1 << 30)}
    if kw["datatype"] in lo_hi:
        lo, hi = lo_hi[kw["datatype"]]
        slope = kw.get("scl", (0.0, 0.0))[0]
        if kw["datatype"] == 2 and slope > 1.0:
            hi = int(255 / slope)  # a scaled uint8 read must stay in range (else C++ leaves it undefined)
        vol = np.clip(vol, lo, hi)
    nifti_files.write(path, vol, **kw)
    return path


@pytest.mark.parametrize("name", sorted(CASES))
def test_read_matches_reference(tmp_path, ref, name):
    tipl_amd = _lib()
    vol = _vol((2, 5, 6, 7), seed=len(name))
    path = _write_case(tmp_path, name, vol)
    for volume in (0, 1):
        got, info = tipl_amd.nifti_read(path, volume)
        want, vs, srow, mni = ref.nifti_read(path, volume)
        assert got.shape == want.shape, name
        assert np.array_equal(_bits(got), _bits(want)), name
        assert np.array_equal(_bits(info["vs"]), _bits(vs)), name
        assert np.array_equal(_bits(info["srow"]), _bits(srow)), name
        assert info["is_mni"] == mni
    if CASES[name]["datatype"] in (2, 256):
        got, _ = tipl_amd.nifti_read(path, 0, np.uint8)
        want, _, _, _ = ref.nifti_read(path, 0, np.uint8)
        assert np.array_equal(got, want)


def test_written_files_read_by_reference(tmp_path, ref):
    tipl_amd = _lib()
    vol = _vol((9, 8, 7), seed=3).astype(np.float32) / 7
    srow = np.array([-1.2, 0, 0, 40, 0, -1.1, 0, 50, 0, 0, 1.3, -20], np.float32)
    for name in ("a.nii", "b.nii.gz"):
        tipl_amd.nifti_write(tmp_path / name, vol, (1.2, 1.1, 1.3), srow)
        got, vs, s, _ = ref.nifti_read(tmp_path / name)
        assert np.array_equal(_bits(got), _bits(vol))
        assert np.array_equal(vs, np.array([1.2, 1.1, 1.3], np.float32))
        assert np.array_equal(_bits(s), _bits(srow))
        # and the reference's own writer gives a file the library reads the same way
        ref.nifti_write(tmp_path / ("r" + name), vol, (1.2, 1.1, 1.3), srow)
        a, ia = tipl_amd.nifti_read(tmp_path / ("r" + name))
        b, ib = tipl_amd.nifti_read(tmp_path / name)
        assert np.array_equal(_bits(a), _bits(b)) and np.array_equal(_bits(a), _bits(vol))
        assert np.array_equal(ia["srow"], ib["srow"])


def test_orientation_closed_forms(tmp_path):
    """RAS (positive diagonal) flips x and y into LPS; the translation moves
    to the opposite corner; a swapped-axis file transposes x and y."""
    tipl_amd = _lib()
    vol = _vol((4, 5, 6), seed=1).astype(np.float32)
    nifti_files.write(tmp_path / "ras.nii", vol, datatype=16, vs=(2, 3, 4),
                      srow=[2, 0, 0, -10, 0, 3, 0, -20, 0, 0, 4, -30])
    got, info = tipl_amd.nifti_read(tmp_path / "ras.nii")
    assert np.array_equal(got, vol[:, ::-1, ::-1])
    assert info["flip_swap"] == [0, 1]
    assert np.allclose(info["srow"], [-2, 0, 0, -10 + 2 * 5, 0, -3, 0, -20 + 3 * 4, 0, 0, 4, -30])
    nifti_files.write(tmp_path / "sw.nii", vol, datatype=16,
                      srow=[0, -1, 0, 0, -1, 0, 0, 0, 0, 0, 1, 0])
    got, info = tipl_amd.nifti_read(tmp_path / "sw.nii")
    assert info["flip_swap"] == [3]
    assert got.shape == (4, 6, 5) and np.array_equal(got, vol.transpose(0, 2, 1))


def test_encodings_read_the_same(tmp_path):
    """Byte order, gzip, NIfTI-2 and Analyze pairs change nothing read; scl is
    slope then intercept in the output type; a second volume follows the first."""
    tipl_amd = _lib()
    vol = _vol((2, 5, 6, 7), seed=2)
    base = dict(datatype=4, srow=[-1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1, 0])
    paths = []
    for i, extra in enumerate([{}, {"endian": ">"}, {"gz": True}, {"version": 2},
                               {"version": 2, "endian": ">", "gz": True}]):
        p = tmp_path / (f"v{i}.nii" + (".gz" if extra.get("gz") else ""))
        nifti_files.write(p, vol, **base, **extra)
        paths.append(p)
    nifti_files.write(tmp_path / "an.hdr", vol, version=0, **base)
    paths.append(tmp_path / "an.hdr")
    first = [tipl_amd.nifti_read(p, 0)[0] for p in paths]
    second = [tipl_amd.nifti_read(p, 1)[0] for p in paths]
    for a, b in zip(first, second):
        assert np.array_equal(a, first[0]) and np.array_equal(b, second[0])
    assert np.array_equal(first[0], vol[0].astype(np.float32))
    assert np.array_equal(second[0], vol[1].astype(np.float32))
    nifti_files.write(tmp_path / "s.nii", vol[0], datatype=4, scl=(0.5, 10.0), srow=base["srow"])
    got, _ = tipl_amd.nifti_read(tmp_path / "s.nii")
    assert np.array_equal(got, vol[0].astype(np.float32) * np.float32(0.5) + np.float32(10.0))


def test_errors(tmp_path):
    tipl_amd = _lib()
    with pytest.raises(tipl_amd.TiplHipError):
        tipl_amd.nifti_read(tmp_path / "missing.nii")
    (tmp_path / "bad.nii").write_bytes(b"\0" * 400)
    with pytest.raises(tipl_amd.TiplHipError):
        tipl_amd.nifti_read(tmp_path / "bad.nii")
    nifti_files.write(tmp_path / "short.nii", _vol((5, 6, 7)), datatype=16)
    data = (tmp_path / "short.nii").read_bytes()
    (tmp_path / "short.nii").write_bytes(data[:-100])
    with pytest.raises(tipl_amd.TiplHipError):
        tipl_amd.nifti_read(tmp_path / "short.nii")
    # malformed headers are refused, never read past the file
    import struct
    nifti_files.write(tmp_path / "ok.nii", _vol((2, 5, 6, 7)), datatype=16)
    good = bytearray((tmp_path / "ok.nii").read_bytes())
    for name, off, fmt, val in (("bitpix", 72, "<h", 8),           # float32 data, 8-bit bitpix
                                ("neg_offset", 108, "<f", -4096.0),
                                ("nan_offset", 108, "<f", float("nan")),
                                ("huge_offset", 108, "<f", 3.0e38)):
        bad = bytearray(good)
        struct.pack_into(fmt, bad, off, val)
        (tmp_path / (name + ".nii")).write_bytes(bytes(bad))
        for volume in (0, 1):
            with pytest.raises(tipl_amd.TiplHipError):
                tipl_amd.nifti_read(tmp_path / (name + ".nii"), volume)
    big = bytearray(good)  # dims far beyond the voxels the file holds
    struct.pack_into("<8h", big, 40, 4, 512, 512, 512, 2, 1, 1, 1)
    (tmp_path / "big.nii").write_bytes(bytes(big))
    for volume in (0, 1):
        with pytest.raises(tipl_amd.TiplHipError):
            tipl_amd.nifti_read(tmp_path / "big.nii", volume, np.uint8)


def test_cpp_mirror_reads_and_writes(tmp_path, ref):
    """tipl::io::gz_nifti of the C++ mirror (tests/cpp/reg_cli) on a swapped,
    gzip-compressed NIfTI-2 file: voxels, shape, voxel size, transformation
    equal the reference reader's; the file it writes back reads the same."""
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp",
                       "reg_cli")
    if not os.path.exists(cli):
        pytest.skip("tests/cpp/reg_cli is not built")
    vol = _vol((2, 5, 6, 7), seed=9)
    path = _write_case(tmp_path, "nifti2_gz", vol)
    out = tmp_path / "o"
    r = subprocess.run([cli, "nifti", str(path), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    want, vs, srow, mni = ref.nifti_read(path)
    txt = (tmp_path / "o.txt").read_text().split("\n")
    shape = tuple(int(v) for v in txt[0].split())
    assert shape == want.shape[::-1]
    got = np.fromfile(tmp_path / "o.f32", np.float32).reshape(want.shape)
    assert np.array_equal(_bits(got), _bits(want))
    assert np.array_equal(np.array(txt[1].split(), np.float32), vs)
    assert np.array_equal(np.array(txt[2].split(), np.float32), srow)
    back, vs2, srow2, _ = ref.nifti_read(tmp_path / "o.nii.gz")
    assert np.array_equal(_bits(back), _bits(want)) and np.array_equal(srow2, srow)

"""The MATLAB Level 5 velocity export (qlp_export_*) read back by scipy's
independent MAT-file reader, plus the header bytes the reference writes
(exporter/MatlabLevel5Header.java :62-108).  Host code only: runs on the CPU."""
import numpy as np
import pytest
import scipy.io

import qlpiv

DATE = "Thu Dec 03 12:00:00 UTC 2026"


def grid(rng, nI, nJ):
    return (rng.normal(0, 3, (nI, nJ)).astype(np.float32),
            rng.normal(0, 3, (nI, nJ)).astype(np.float32))


def test_multi_frame_export_reads_back(tmp_path):
    rng = np.random.default_rng(1)
    path = tmp_path / "velocities.mat"
    maps = {}
    with qlpiv.VelocityExporter(path, (768, 1024), (64, 64), (16, 16), 0.5, total_maps=3,
                                date_text=DATE) as ex:
        for f in range(3):
            for step, (N, nI, nJ) in enumerate(((64, 23, 31), (32, 47, 63), (16, 95, 127))):
                u, v = grid(rng, nI, nJ)
                maps[(N, f)] = (u, v)
                ex.export(f, u, v, (N, N), step, margins=(1, 2, 3, 4))
    raw = path.read_bytes()
    title = f"MATLAB 5.0 - QuickLabPIV-ng Velocities exported on: {DATE}".encode()
    assert raw[:len(title)] == title and raw[len(title):116] == b"\0" * (116 - len(title))
    assert raw[116:124] == b"\0" * 8 and raw[124:126] == b"\x00\x01" and raw[126:128] == b"IM"
    m = scipy.io.loadmat(path, squeeze_me=False, mat_dtype=False)
    assert str(m["exportDate"][0]) == DATE
    p = m["parameters"][0, 0]
    assert p["imageWidth"][0, 0] == 1024 and p["imageHeight"][0, 0] == 768
    assert p["imageWidth"].dtype == np.int32
    assert p["iaStartI"][0, 0] == 64 and p["iaEndJ"][0, 0] == 16
    assert p["overlapFactor"].dtype == np.float32 and p["overlapFactor"][0, 0] == np.float32(0.5)
    assert p["totalNumberOfVelocityMaps"][0, 0] == 3 and p["multiVolumeLastVolume"][0, 0] == 1
    assert list(p.dtype.names) == ["imageWidth", "imageHeight", "iaStartI", "iaEndI", "iaStartJ",
                                   "iaEndJ", "overlapFactor", "hasSwappedUVVectors",
                                   "markInvalidAsNaN", "totalNumberOfVelocityMaps",
                                   "multiVolumeExport", "multiVolumeLastVolume"]
    # variables in creation order, one struct per IA size
    assert [k for k in m if not k.startswith("__")] == ["exportDate", "parameters", "velocities64x64",
                                                        "velocities32x32", "velocities16x16"]
    for step, N in enumerate((64, 32, 16)):
        s = m[f"velocities{N}x{N}"][0, 0]
        assert list(s.dtype.names) == ["iaHeight", "iaWidth", "adaptiveStep", "margins", "u", "v"]
        assert s["iaHeight"][0, 0] == N and s["adaptiveStep"][0, 0] == step
        mg = s["margins"][0, 0]
        assert (mg["top"][0, 0], mg["left"][0, 0], mg["bottom"][0, 0], mg["right"][0, 0]) == (1, 2, 3, 4)
        assert s["u"].dtype == np.float32 and s["u"].shape[2] == 3
        for f in range(3):
            u, v = maps[(N, f)]
            assert np.array_equal(s["u"][:, :, f], u) and np.array_equal(s["v"][:, :, f], v)


def test_invalid_vectors_as_nan_and_multi_volume(tmp_path):
    rng = np.random.default_rng(2)
    path = tmp_path / "part1.mat"
    u, v = grid(rng, 7, 9)
    inv = rng.random((7, 9)) < 0.3
    with qlpiv.VelocityExporter(path, (256, 320), (32, 32), (32, 32), 0.5, total_maps=2,
                                mark_invalid_nan=True, swap_uv=True, multi_volume=True,
                                next_filename="part2.mat", date_text=DATE) as ex:
        ex.export(1, u, v, (32, 32), 0, invalid=inv)
    m = scipy.io.loadmat(path)
    p = m["parameters"][0, 0]
    assert p["markInvalidAsNaN"][0, 0] == 1 and p["hasSwappedUVVectors"][0, 0] == 1
    assert p["multiVolumeExport"][0, 0] == 1 and p["multiVolumeLastVolume"][0, 0] == 0
    assert str(p["multiVolumeNextFilename"][0]) == "part2.mat"
    s = m["velocities32x32"][0, 0]
    got_u = s["u"][:, :, 1]
    assert np.array_equal(np.isnan(got_u), inv)
    assert np.array_equal(got_u[~inv], u[~inv]) and np.array_equal(s["v"][:, :, 1][~inv], v[~inv])
    assert np.all(s["u"][:, :, 0] == 0)  # a frame never exported stays zero


def test_export_errors(tmp_path):
    with pytest.raises(qlpiv.QlpError):
        qlpiv.VelocityExporter(tmp_path / "x.mat", (10, 10), (16, 16), (16, 16), 0.5, total_maps=0)
    with pytest.raises(qlpiv.QlpError):
        qlpiv.VelocityExporter(tmp_path / "no" / "such" / "dir.mat", (10, 10), (16, 16), (16, 16), 0.5)
    ex = qlpiv.VelocityExporter(tmp_path / "y.mat", (10, 10), (16, 16), (16, 16), 0.5, total_maps=1)
    u = np.zeros((3, 4), np.float32)
    with pytest.raises(qlpiv.QlpError):
        ex.export(1, u, u, (16, 16), 0)  # frame beyond total_maps
    ex.export(0, u, u, (16, 16), 0)
    with pytest.raises(qlpiv.QlpError):
        ex.export(0, np.zeros((2, 4), np.float32), np.zeros((2, 4), np.float32), (16, 16), 0)
    ex.close()


@pytest.mark.parametrize("multi", [False, True])
def test_whole_results_export(tmp_path, multi):
    """exportDataToFile(PIVResults) :185-238: one struct per velocity map with
    iaHeight, iaWidth, multiVolumeExport (and multiVolumeStartOffset for a
    multi-volume file), margins and u, v of nI x nJ x maps, frames at their map
    indices, values as they are (NaN kept, no swap); maps never given are 0."""
    rng = np.random.default_rng(7)
    path = tmp_path / "results.mat"
    kw = dict(multi_volume=True, next_filename="next.mat") if multi else {}
    u32, v32 = rng.normal(0, 2, (2, 11, 13)).astype(np.float32), rng.normal(0, 2, (2, 11, 13)).astype(np.float32)
    u32[1, 3, 4] = np.nan
    u16, v16 = rng.normal(0, 2, (3, 23, 27)).astype(np.float32), rng.normal(0, 2, (3, 23, 27)).astype(np.float32)
    with qlpiv.VelocityExporter(path, (384, 448), (32, 32), (16, 16), 0.5, total_maps=4,
                                date_text=DATE, **kw) as ex:
        ex.export_results([0, 2], u32, v32, (32, 32), 4, margins=(5, 6, 7, 8), start_offset=40)
        ex.export_results([0, 1, 3], u16, v16, (16, 16), 4, margins=(1, 2, 3, 4), start_offset=40)
    m = scipy.io.loadmat(path)
    names = [k for k in m if k.startswith("velocities")]
    assert names == ["velocities32x32", "velocities16x16"]
    for name, (u, v, frames, marg) in {"velocities32x32": (u32, v32, [0, 2], (5, 6, 7, 8)),
                                        "velocities16x16": (u16, v16, [0, 1, 3], (1, 2, 3, 4))}.items():
        s = m[name][0, 0]
        fields = list(s.dtype.names)
        want = ["iaHeight", "iaWidth", "multiVolumeExport"] + (["multiVolumeStartOffset"] if multi else []) + \
               ["margins", "u", "v"]
        assert fields == want
        assert s["multiVolumeExport"][0, 0] == (1 if multi else 0)
        if multi:
            assert s["multiVolumeStartOffset"][0, 0] == 40
        mg = s["margins"][0, 0]
        assert [int(mg[k][0, 0]) for k in ("top", "left", "bottom", "right")] == list(marg)
        assert s["u"].shape == u.shape[1:] + (4,) and s["u"].dtype == np.float32
        for k, fr in enumerate(frames):
            assert np.array_equal(s["u"][:, :, fr], u[k], equal_nan=True)
            assert np.array_equal(s["v"][:, :, fr], v[k])
        for fr in set(range(4)) - set(frames):
            assert not s["u"][:, :, fr].any()

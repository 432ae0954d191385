"""HDF5 files of the reference's formats, read and written without h5py.

The reader is pinned by a file the HDF5 library wrote: the reference's own
phantom input real_left_ventricle_contraction.h5 (phantom_data/, copied to
tests/golden/ unchanged; phantom_scripts/phantom_common.py:5-25 reads its
"times" and "factors").  The writer is pinned by that file too: written with
the same two datasets, its superblock, root group header and group B-tree node
header are byte for byte the library's (only the end-of-file address differs,
the files' sizes being different).  Round trips cover every element type,
rank 0-3, empty datasets and groups of more than one symbol-table node.  The
loaders check the names, shapes and errors of HDFConvenience.cpp:45-196 and
drive the FP64 oracle to the same frame as the in-memory arrays.
"""
import os

import numpy as np
import pytest

from backends import Oracle
from openbcsim_amd import h5, hdf
from openbcsim_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
LV = os.path.join(HERE, "golden", "real_left_ventricle_contraction.h5")


def test_reads_a_file_the_hdf5_library_wrote():
    d = h5.read(LV)
    assert sorted(d) == ["factors", "times"]
    t, f = d["times"], d["factors"]
    assert t.dtype == np.float32 and f.dtype == np.float32 and t.shape == f.shape == (98,)
    # a sampled time axis from 0 and a smooth contraction factor in (0, 1]: a
    # misread offset, length or byte order gives garbage, not these
    assert t[0] == 0.0 and np.all(np.diff(t) > 0) and abs(float(t[-1]) - 1.0058) < 1e-3
    assert np.all(np.isfinite(f)) and f.min() == np.float32(0.75) and f.max() == np.float32(1.0)
    assert np.max(np.abs(np.diff(f))) < 0.05


def test_writer_matches_the_library_layout(tmp_path):
    d = h5.read(LV)
    out = str(tmp_path / "w.h5")
    h5.write(out, {"times": d["times"], "factors": d["factors"]})
    lib = open(LV, "rb").read()
    mine = open(out, "rb").read()
    # superblock and root-group symbol-table entry (96 bytes), root object
    # header (40): identical except the end-of-file address at bytes 40..47
    assert lib[:40] == mine[:40] and lib[48:136] == mine[48:136]
    assert int.from_bytes(mine[40:48], "little") == len(mine)
    # group B-tree node at 136: signature, type, level, entries, siblings, key 0
    assert lib[136:168] == mine[136:168]
    back = h5.read(out)
    assert all(np.array_equal(back[k], d[k]) for k in d)


def test_round_trip_types_and_shapes(tmp_path):
    rng = np.random.default_rng(0)
    data = {"scalar_int": np.int64(3), "scalar_f": np.float32(0.125), "empty": np.zeros((0, 4), np.float32),
            "f16": np.array([1.5, -2.0], np.float16), "f32": rng.random((7, 4)).astype(np.float32),
            "f64": rng.random((3, 5, 3)), "big_endian": np.arange(12, dtype=">f4").reshape(3, 4),
            "i8": np.array([-128, 0, 127], np.int8), "u16": np.array([0, 65535], np.uint16),
            "i32": np.arange(-3, 3, dtype=np.int32).reshape(2, 3), "u64": np.array([2 ** 63 + 5], np.uint64)}
    for k in range(19):  # more names than one symbol-table node holds (8)
        data[f"n{k:02d}"] = np.full(2, k, np.int16)
    path = str(tmp_path / "r.h5")
    h5.write(path, data)
    back = h5.read(path)
    assert sorted(back) == sorted(data)
    for k, v in data.items():
        v = np.asarray(v)
        assert back[k].shape == v.shape and back[k].dtype == v.dtype.newbyteorder("="), k
        assert np.array_equal(back[k], v), k


def test_reader_and_writer_errors(tmp_path):
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not an hdf5 file" * 8)
    with pytest.raises(h5.H5Error, match="signature"):
        h5.read(str(bad))
    with pytest.raises(h5.H5Error, match="name"):
        h5.write(str(tmp_path / "x.h5"), {"a/b": np.zeros(2)})
    with pytest.raises(h5.H5Error, match="type"):
        h5.write(str(tmp_path / "x.h5"), {"s": np.array(["text"])})


def test_phantom_file_round_trips_and_checks(tmp_path):
    rng = np.random.default_rng(1)
    pts = rng.random((50, 4)).astype(np.float32)
    p = str(tmp_path / "fixed.h5")
    hdf.save_fixed_scatterers(p, pts)
    assert np.array_equal(hdf.load_fixed_scatterers(p), pts)
    h5.write(p, {"data": rng.random((5, 3))})
    with pytest.raises(RuntimeError, match="four"):
        hdf.load_fixed_scatterers(p)

    cp = rng.random((6, 10, 3))  # float64, as the phantom scripts write it
    s = str(tmp_path / "spline.h5")
    h5.write(s, {"spline_degree": np.int64(3), "control_points": cp, "amplitudes": np.ones(6, np.float32),
                 "knot_vector": np.linspace(0, 1, 14).astype(np.float32)})
    sp = hdf.load_spline_scatterers(s)
    assert sp.spline_degree == 3 and sp.control_points.dtype == np.float32
    assert np.array_equal(sp.control_points, cp.astype(np.float32)) and sp.knot_vector.shape == (14,)

    lut = rng.random((8, 4, 5)).astype(np.float32) * 3.0
    b = str(tmp_path / "lut.h5")
    hdf.save_beam_profile(b, lut, (0.0, 0.12), (-0.01, 0.01), (-0.02, 0.02))
    bp = hdf.load_beam_profile(b)
    assert np.allclose(bp.samples, lut / lut.max(), rtol=0, atol=0) and bp.samples.max() == 1.0
    assert bp.rad_extent == (0.0, pytest.approx(0.12)) and bp.lat_extent[0] == pytest.approx(-0.01)
    h5.write(b, {"beam_profile": lut, "rad_extent": np.zeros(3), "lat_extent": np.zeros(2),
                 "ele_extent": np.zeros(2)})
    with pytest.raises(RuntimeError, match="rad_extent"):
        hdf.load_beam_profile(b)

    o, d, l, t, length = W.sector_scan(16, np.radians(60.0), 0.12)
    sc = str(tmp_path / "scan.h5")
    hdf.save_scan_sequence(sc, o, d, l, t, length)
    seq = hdf.load_scan_sequence(sc)
    assert np.array_equal(seq.origins, o) and np.array_equal(seq.lateral_dirs, l)
    assert seq.line_length == pytest.approx(length)

    e = str(tmp_path / "exc.h5")
    hdf.save_excitation(e, np.arange(5, dtype=np.float32), 2, 50e6)
    ex = hdf.load_excitation(e)
    assert ex.center_index == 2 and ex.sampling_frequency == 50e6 and ex.samples.shape == (5,)


def _config(sim, wl):
    for k, v in wl.params.items():
        sim.set_parameter(k, str(v))
    sim.set_excitation(wl.excitation, wl.center_index, wl.fs, wl.demod_freq)


def test_files_drive_the_oracle_to_the_same_frame(tmp_path):
    """A spline phantom with its scan, and a fixed phantom with a LUT profile,
    written to files and loaded back give the oracle the frame of the in-memory
    arrays, bit for bit (the LUT as the file loader normalises it: divided by
    its maximum, HDFConvenience.cpp:179-193)."""
    wl = W.spline_mmode(300, 12, seed=7, num_cp=6, decimation=2)
    direct = Oracle()
    wl.apply(direct)
    sp = wl.splines[0]
    spath, cpath = str(tmp_path / "s.h5"), str(tmp_path / "c.h5")
    hdf.save_spline_scatterers(spath, sp.control_points, sp.amplitudes, sp.knots, sp.degree)
    o, d, l, t, length = wl.scan
    hdf.save_scan_sequence(cpath, o, d, l, t, length)
    via = Oracle()
    _config(via, wl)
    hdf.set_scan_sequence_from_hdf(via, cpath)
    hdf.add_spline_scatterers_from_hdf(via, spath)
    via.set_analytical_beam_profile(wl.profile[1], wl.profile[2])
    assert np.array_equal(via.simulate_frame(), direct.simulate_frame())

    vol = W.volume_lut(2000, n_az=3, n_el=3, seed=8, lut_shape=(32, 8, 8))
    (r0, r1, l0, l1, e0, e1), table = vol.profile[1], vol.profile[2]
    fpath, bpath = str(tmp_path / "f.h5"), str(tmp_path / "b.h5")
    hdf.save_fixed_scatterers(fpath, vol.fixed[0])
    hdf.save_beam_profile(bpath, table * 2.5, (r0, r1), (l0, l1), (e0, e1))
    direct = Oracle()
    _config(direct, vol)
    o, d, l, t, length = vol.scan
    direct.set_scan_sequence(o, d, length, l, t)
    direct.add_fixed_scatterers(vol.fixed[0])
    direct.set_lut_beam_profile(r0, r1, l0, l1, e0, e1, (table * 2.5) / (table * 2.5).max())
    via = Oracle()
    _config(via, vol)
    via.set_scan_sequence(o, d, length, l, t)
    hdf.set_fixed_scatterers_from_hdf(via, fpath)
    hdf.set_beam_profile_from_hdf(via, bpath)
    frame = via.simulate_frame()
    assert np.abs(frame).max() > 0 and np.array_equal(frame, direct.simulate_frame())

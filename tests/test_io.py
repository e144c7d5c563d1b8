"""State and MPO files (imps_io.cpp through the C ABI): bit-exact round trips and refusal of
missing, foreign, truncated and corrupted files.  Host-only: no GPU needed."""
import ctypes
import os

import numpy as np
import pytest

import infinitempsalgorithms_amd as imps
from infinitempsalgorithms_amd import _lib


def _state(oracle, D=5, ds=(3, 3), seed=4):
    psi = oracle.randomimps(D, list(ds), seed)
    return imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)


def test_state_file_round_trip_is_bit_exact(oracle, tmp_path):
    st = _state(oracle)
    path = tmp_path / "psi.imps"
    imps.save_state(st, path)
    assert not os.path.exists(str(path) + ".tmp")
    n, D, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L = _lib.load()
    assert L.imps_state_file_info(os.fsencode(path), ctypes.byref(n), ctypes.byref(D), ctypes.byref(d)) == 0
    assert (n.value, D.value, d.value) == (2, 5, 3)
    back = imps.load_state(path)
    for xs, ys in zip((st.AL, st.AR, st.C, st.AC), (back.AL, back.AR, back.C, back.AC)):
        for x, y in zip(xs, ys):
            assert np.array_equal(x, y)
    # header 32 bytes, payload, 8-byte hash
    assert os.path.getsize(path) == 32 + 8 * 2 * (3 * 5 * 3 * 5 + 25) + 8
    # overwriting keeps one complete file
    imps.save_state(_state(oracle, seed=5), path)
    assert not np.array_equal(imps.load_state(path).AL[0], st.AL[0])


def test_mpo_file_round_trip(tmp_path):
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=2)
    path = tmp_path / "h.impo"
    imps.save_mpo(H, path)
    back = imps.load_mpo(path)
    assert isinstance(back, imps.SparseIMPO) and len(back) == 2
    for x, y in zip(H.Ws, back.Ws):
        assert np.array_equal(x, y)
    dense = imps.load_mpo(path, dense=True)
    assert isinstance(dense, imps.DenseIMPO) and np.array_equal(dense[1], H.Ws[1])


def test_bad_files_are_refused_and_outputs_untouched(oracle, tmp_path):
    st = _state(oracle)
    path = tmp_path / "psi.imps"
    imps.save_state(st, path)
    raw = bytearray(path.read_bytes())
    L = _lib.load()

    def read_raw(p, N=2, D=5, d=3):
        outs = [[np.full((D, d, D), 7.0, order="F") for _ in range(N)] for _ in range(2)]
        C = [np.full((D, D), 7.0, order="F") for _ in range(N)]
        AC = [np.full((D, d, D), 7.0, order="F") for _ in range(N)]
        keep = [_lib.ptr_array(x) for x in (outs[0], outs[1], C, AC)]
        st_ = L.imps_read_state_file(os.fsencode(p), N, D, d, *(k[0] for k in keep))
        untouched = all(np.all(x == 7.0) for xs in (outs[0], outs[1], C, AC) for x in xs)
        return st_, untouched, L.imps_last_error(None).decode()

    # one flipped payload byte: checksum
    bad = bytearray(raw)
    bad[32 + 100] ^= 0x10
    p = tmp_path / "flip.imps"
    p.write_bytes(bytes(bad))
    st_, untouched, msg = read_raw(p)
    assert st_ == _lib.IMPS_E_IO and untouched and "checksum" in msg
    # truncated payload, truncated hash, trailing bytes
    for name, data in (("trunc", raw[:200]), ("nohash", raw[:-4]), ("extra", raw + b"\0")):
        p = tmp_path / f"{name}.imps"
        p.write_bytes(bytes(data))
        st_, untouched, msg = read_raw(p)
        assert st_ == _lib.IMPS_E_IO and untouched, (name, msg)
    # another file kind, a missing file, a short header
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=2)
    imps.save_mpo(H, tmp_path / "h.impo")
    for p, word in ((tmp_path / "h.impo", "not a state file"), (tmp_path / "missing.imps", "cannot open")):
        st_, untouched, msg = read_raw(p)
        assert st_ == _lib.IMPS_E_IO and untouched and word in msg
    (tmp_path / "short.imps").write_bytes(raw[:10])
    with pytest.raises(_lib.ImpsError, match="truncated header"):
        imps.load_state(tmp_path / "short.imps")
    # requested dimensions differ from the file's
    st_, untouched, msg = read_raw(path, D=4)
    assert st_ == _lib.IMPS_E_ARG and untouched
    # a write into a missing directory fails cleanly
    with pytest.raises(_lib.ImpsError, match="cannot open"):
        imps.save_state(st, tmp_path / "no" / "such" / "dir.imps")
    with pytest.raises(_lib.ImpsError, match="checksum"):
        imps.load_state(tmp_path / "flip.imps")


def test_headers_with_wrong_or_absurd_dimensions_are_refused(oracle, tmp_path):
    # the reader compares the file size with the size the header announces before allocating,
    # so a damaged header cannot make it allocate or read past the payload
    import struct
    st = _state(oracle)
    path = tmp_path / "psi.imps"
    imps.save_state(st, path)
    raw = bytearray(path.read_bytes())
    for nsites, D, d in ((1_000_000_000, 100_000, 4), (2, 6, 3), (3, 5, 3), (2, 5, 2)):
        bad = bytearray(raw)
        struct.pack_into("<iii", bad, 16, nsites, D, d)
        p = tmp_path / f"hdr_{nsites}_{D}_{d}.imps"
        p.write_bytes(bytes(bad))
        with pytest.raises(_lib.ImpsError, match="not plausible|does not match"):
            imps.load_state(p)
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=2)
    imps.save_mpo(H, tmp_path / "h.impo")
    bad = bytearray((tmp_path / "h.impo").read_bytes())
    struct.pack_into("<i", bad, 16, 2_000_000_000)
    (tmp_path / "hbad.impo").write_bytes(bytes(bad))
    with pytest.raises(_lib.ImpsError, match="not plausible|does not match"):
        imps.load_mpo(tmp_path / "hbad.impo")

"""The library's frame formatter (ndg_format_frame, used by ndg_frame_write) against the
reference's write_model edit descriptors (src/read_input.f90:403-421) as the Fortran
compiler formats them: tests/fortran/frames_ref.f90 writes the same values with
(A4,x,7EN14.4) and (i7,x,i2,3f9.5) / (2f12.4) / (i7).  Bar: identical bytes.

The values include exact decimal ties (k/64 at 5 decimals, 1.03125 at 4), negative zeros,
negatives that round to zero, values that round up to the next power of ten or of 1000
(999.99996, 99.99996), exponents beyond 99 (the 'E'-less form), fields that overflow, and
random values over 30 decades.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "ndyn_omp_amd", "libndyn_gpu.so")
SRC = os.path.join(ROOT, "tests", "fortran", "frames_ref.f90")
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

pytestmark = [pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built"),
              pytest.mark.skipif(FC is None, reason="no Fortran compiler")]

XYZ, DUMP = 1, 2


def values(rng, n):
    special = np.array([0.0, -0.0, 1e-7, -1e-7, 0.015625, -0.015625, 0.046875, 1.03125, -1.09375,
                        999.99996, -999.99996, 99.99996, 9.99996, 0.99999996, 0.999999996e-3,
                        1e-100, -1e-100, 1e100, 1.5e300, 1e-300, 123456.0, -9.87654321e-3,
                        0.49999999, -0.5, 1e6, 12345678.9, 0.1, 1e-3, 1e3, 1e-30])
    r = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-15, 15, n)
    r[: len(special)] = special[: min(n, len(special))]
    return r


def test_formatter_matches_fortran_edit_descriptors(tmp_path):
    rng = np.random.default_rng(7)
    np_ = 5000
    names = ["Cu", "Cu-1", "LongName"]
    typ = rng.integers(1, 4, np_).astype(np.int32)
    pos = rng.uniform(-0.5, 0.5, (np_, 3))
    pos[:64, 0] = (np.arange(64) - 32) / 64.0            # exact ties at 5 decimals
    pos[64:70, 1] = [-0.0, -1e-7, 1e-7, -0.499999, 0.4999999, -0.000004]
    pos[70:75, 2] = [12.5, -99.999995, 999.99999, 1e4, -1e5]  # fields that overflow f9.5
    vel = values(rng, 3 * np_).reshape(np_, 3)
    pene = values(rng, np_)
    box = np.array([87.096, 43.548, 123456789.0])
    step = 4000

    with open(tmp_path / "frame_in.bin", "wb") as f:
        np.array([np_, len(names), step], np.int32).tofile(f)
        f.write("".join(n.ljust(8) for n in names).encode())
        typ.tofile(f)
        pos.tofile(f)          # pos(3,np) column-major == (np,3) row-major
        vel.tofile(f)
        pene.tofile(f)
        box.tofile(f)
    exe = tmp_path / "frames_ref"
    subprocess.run([FC, "-O0", "-o", str(exe), SRC], check=True, capture_output=True)
    subprocess.run([str(exe)], cwd=tmp_path, check=True, capture_output=True)

    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.ndg_format_frame.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.POINTER(C.c_int), C.c_char_p,
                                   C.c_int, dp, C.c_int, C.c_char_p, C.c_size_t,
                                   C.POINTER(C.c_size_t)]
    nm = "".join(n.ljust(8) for n in names).encode()

    def fmt(kind):
        P = lambda a: np.ascontiguousarray(a).ctypes.data_as(dp)
        n = C.c_size_t(0)
        assert L.ndg_format_frame(kind, np_, P(pos), P(vel), P(pene),
                                  typ.ctypes.data_as(C.POINTER(C.c_int)), nm, 8, P(box), step,
                                  None, 0, C.byref(n)) == 0
        buf = C.create_string_buffer(n.value)
        assert L.ndg_format_frame(kind, np_, P(pos), P(vel), P(pene),
                                  typ.ctypes.data_as(C.POINTER(C.c_int)), nm, 8, P(box), step,
                                  buf, n.value, C.byref(n)) == 0
        return buf.raw[: n.value]

    for kind, ref in ((XYZ, "frame_ref.xyz"), (DUMP, "frame_ref.dump")):
        want = (tmp_path / ref).read_bytes()
        got = fmt(kind)
        if got != want:
            gl, wl = got.decode().splitlines(), want.decode().splitlines()
            bad = [k for k, (a, b) in enumerate(zip(gl, wl)) if a != b][:5]
            pytest.fail(f"{ref}: {len(gl)} vs {len(wl)} lines; first differences: " +
                        "; ".join(f"{k}: {gl[k]!r} != {wl[k]!r}" for k in bad))

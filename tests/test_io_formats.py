"""MRC and STAR formats (include/t2x_io.h) against the published layouts, on the CPU.

MRC: the MRC2014 header (1024 bytes of 4-byte words at fixed offsets; mode 0 int8, 1 int16,
2 float32, 6 uint16, 12 float16; nsymbt bytes of extended header before the data; files of the
other endianness detected by the size of mode/nx as rwMRC.h does) is built here byte by byte with
numpy, independently of the library, and read through it; files the library writes are parsed
here at the same offsets. STAR: RELION 3.1 particle files (data_optics + data_particles loops,
"_label #n" lines, comments, quoted values) and list blocks are parsed, converted to E-step
inputs as getMetaAndImageDataSubset does (Angstrom offsets / pixel size, optics-group CTF values,
defaults for absent columns) and written back with the run_itNNN_data.star columns.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import t2x_py as T


def mrc_bytes(data, mode, angpix=1.5, nsymbt=0, big_endian=False):
    """An MRC2014 file image for data[nz][ny][nx] in the given mode (header per the format spec)."""
    nz, ny, nx = data.shape
    e = ">" if big_endian else "<"
    h = bytearray(1024)
    struct.pack_into(e + "4i", h, 0, nx, ny, nz, mode)
    struct.pack_into(e + "3i", h, 28, nx, ny, nz)
    struct.pack_into(e + "6f", h, 40, angpix * nx, angpix * ny, angpix * nz, 90, 90, 90)
    struct.pack_into(e + "3i", h, 64, 1, 2, 3)
    d64 = data.astype(np.float64)
    struct.pack_into(e + "3f", h, 76, float(d64.min()), float(d64.max()), float(d64.mean()))
    struct.pack_into(e + "2i", h, 88, 0, nsymbt)
    h[208:212] = b"MAP "
    h[212:216] = bytes([17, 17, 0, 0]) if big_endian else bytes([68, 65, 0, 0])
    struct.pack_into(e + "f", h, 216, float(d64.std()))
    dt = {0: "i1", 1: "i2", 2: "f4", 6: "u2", 12: "f2"}[mode]
    body = data.astype((">" if big_endian else "<") + dt).tobytes()
    return bytes(h) + bytes(nsymbt) + body


def read_sections(lib, path, first, count, nx, ny):
    out = np.zeros((count, ny, nx), np.float32)
    T.check(lib.t2x_mrc_read_sections(path.encode(), first, count, T.fptr(out)), "read")
    return out


@pytest.mark.parametrize("mode", [0, 1, 2, 6, 12])
@pytest.mark.parametrize("big_endian", [False, True])
def test_mrc_read_modes(tmp_path, mode, big_endian):
    lib = T.load()
    rng = np.random.default_rng(mode + 10 * big_endian)
    lim = {0: 127, 1: 30000, 2: 1e4, 6: 60000, 12: 1000}[mode]
    lo = 0 if mode == 6 else -lim
    data = rng.uniform(lo, lim, size=(5, 6, 7))
    if mode in (0, 1, 6):
        data = np.round(data)
    data = data.astype({0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}[mode])
    p = tmp_path / f"m{mode}{int(big_endian)}.mrc"
    p.write_bytes(mrc_bytes(data, mode, nsymbt=96 if mode == 2 else 0, big_endian=big_endian))
    info = T.MrcInfo()
    T.check(lib.t2x_mrc_info_read(str(p).encode(), C.byref(info)))
    assert (info.nx, info.ny, info.nz, info.mode, info.swapped) == (7, 6, 5, mode, int(big_endian))
    assert info.data_offset == 1024 + (96 if mode == 2 else 0)
    assert abs(info.angpix - 1.5) < 1e-6
    got = read_sections(lib, str(p), 1, 3, 7, 6)
    np.testing.assert_array_equal(got, data[1:4].astype(np.float32))
    bad = np.zeros((6, 6, 7), np.float32)
    assert lib.t2x_mrc_read_sections(str(p).encode(), 3, 3, T.fptr(bad)) == T.T2X_EINVAL


def test_mrc_write_header_layout(tmp_path):
    lib = T.load()
    rng = np.random.default_rng(3)
    data = rng.normal(2.0, 3.0, size=(4, 9, 8)).astype(np.float32)
    p = tmp_path / "w.mrc"
    T.check(lib.t2x_mrc_write(str(p).encode(), 8, 9, 4, 1.25, T.fptr(np.ascontiguousarray(data))))
    b = p.read_bytes()
    assert len(b) == 1024 + data.nbytes
    nx, ny, nz, mode = struct.unpack_from("<4i", b, 0)
    mx, my, mz = struct.unpack_from("<3i", b, 28)
    cella = struct.unpack_from("<3f", b, 40)
    cellb = struct.unpack_from("<3f", b, 52)
    axes = struct.unpack_from("<3i", b, 64)
    dmin, dmax, dmean = struct.unpack_from("<3f", b, 76)
    ispg, nsymbt = struct.unpack_from("<2i", b, 88)
    rms, = struct.unpack_from("<f", b, 216)
    nlabl, = struct.unpack_from("<i", b, 220)
    assert (nx, ny, nz, mode) == (8, 9, 4, 2) and (mx, my, mz) == (8, 9, 4)
    np.testing.assert_allclose(cella, (10.0, 11.25, 5.0), rtol=1e-6)
    assert cellb == (90.0, 90.0, 90.0) and axes == (1, 2, 3) and nsymbt == 0 and ispg == 0
    assert b[208:212] == b"MAP " and b[212:214] == bytes([68, 65]) and nlabl == 1
    np.testing.assert_allclose((dmin, dmax, dmean, rms), (data.min(), data.max(), data.mean(), data.std()),
                               rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(np.frombuffer(b, "<f4", offset=1024).reshape(4, 9, 8), data)
    np.testing.assert_array_equal(read_sections(lib, str(p), 0, 4, 8, 9), data)


def test_mrc_rejects_fourier_modes(tmp_path):
    lib = T.load()
    p = tmp_path / "c.mrc"
    p.write_bytes(mrc_bytes(np.zeros((1, 2, 2), np.float32), 2)[:12] + struct.pack("<i", 4) +
                  mrc_bytes(np.zeros((1, 2, 2), np.float32), 2)[16:])
    info = T.MrcInfo()
    assert lib.t2x_mrc_info_read(str(p).encode(), C.byref(info)) == T.T2X_EINVAL
    assert b"real-space" in lib.t2x_last_error()


PARTICLES = """
# version 30001

data_optics

loop_
_rlnOpticsGroupName #1
_rlnOpticsGroup #2
_rlnVoltage #3
_rlnSphericalAberration #4
_rlnAmplitudeContrast #5
_rlnImagePixelSize #6
_rlnImageSize #7
_rlnImageDimensionality #8
opticsGroup1 1 300.000000 2.700000 0.100000 1.100000 256 2
"group two" 2 200.000000 2.000000 0.070000 0.880000 256 2

# version 30001

data_particles

loop_
_rlnImageName #1
_rlnDefocusU #2
_rlnDefocusV #3
_rlnDefocusAngle #4
_rlnOpticsGroup #5
_rlnOriginXAngst #6
_rlnOriginYAngst #7
_rlnAngleRot #8
_rlnAngleTilt #9
_rlnAnglePsi #10
_rlnMicrographName #11
000001@Particles/stack_a.mrcs 12000.5 11800.25 33.5 1 2.2 -1.1 10.0 20.0 30.0 mic1.mrc
000002@Particles/stack_a.mrcs 15000.0 14500.0 120.0 2 0.88 0.0 -40.0 95.5 -170.0 "mic 2.mrc"
# a comment line inside the data
000007@Particles/stack_b.mrcs 20000.0 20000.0 0.0 1 0.0 3.3 0.0 0.0 0.0 mic3.mrc
"""


def read_star(lib, path):
    h = C.c_void_p()
    T.check(lib.t2x_star_read(str(path).encode(), C.byref(h)), "star read")
    return h


def test_star_particles_to_estep_inputs(tmp_path):
    lib = T.load()
    p = tmp_path / "particles.star"
    p.write_text(PARTICLES)
    h = read_star(lib, p)
    n = C.c_long()
    T.check(lib.t2x_star_num_rows(h, b"particles", C.byref(n)))
    assert n.value == 3
    T.check(lib.t2x_star_num_rows(h, b"optics", C.byref(n)))
    assert n.value == 2
    pin = (T.ParticleIn * 3)()
    px = np.zeros(3)
    T.check(lib.t2x_star_particles(h, 0, 3, pin, T.dptr(px)))
    np.testing.assert_array_equal(px, [1.1, 0.88, 1.1])
    a, b, c = pin
    assert (a.ctf.voltage_kv, a.ctf.cs_mm, a.ctf.q0) == (300.0, 2.7, 0.1)
    assert (b.ctf.voltage_kv, b.ctf.cs_mm, b.ctf.q0) == (200.0, 2.0, 0.07)
    assert (a.ctf.defocus_u, a.ctf.defocus_v, a.ctf.defocus_angle) == (12000.5, 11800.25, 33.5)
    assert (a.ctf.bfactor, a.ctf.scale, a.ctf.phase_shift) == (0.0, 1.0, 0.0)      # absent columns
    assert abs(a.old_offset_x - 2.0) < 1e-12 and abs(a.old_offset_y + 1.0) < 1e-12  # Angstrom / 1.1
    assert abs(b.old_offset_x - 1.0) < 1e-12 and abs(c.old_offset_y - 3.0) < 1e-12
    assert a.prior_offset_x == 999.0 and a.norm_correction == 1.0
    assert (b.rot, b.tilt, b.psi) == (-40.0, 95.5, -170.0)
    idx = C.c_long()
    path = C.create_string_buffer(256)
    T.check(lib.t2x_star_image_ref(h, 2, C.byref(idx), path, 256))
    assert idx.value == 7 and path.value == b"Particles/stack_b.mrcs"
    s = C.create_string_buffer(64)
    T.check(lib.t2x_star_get_string(h, b"particles", b"rlnMicrographName", 1, s, 64))
    assert s.value == b"mic 2.mrc"
    T.check(lib.t2x_star_get_string(h, b"optics", b"rlnOpticsGroupName", 1, s, 64))
    assert s.value == b"group two"
    v = C.c_double()
    assert lib.t2x_star_get_double(h, b"particles", b"rlnNoSuchLabel", 0, C.byref(v)) == T.T2X_EINVAL
    lib.t2x_star_free(h)


def test_star_store_results_round_trip(tmp_path):
    lib = T.load()
    p = tmp_path / "particles.star"
    p.write_text(PARTICLES)
    h = read_star(lib, p)
    out = (T.ParticleOut * 2)()
    out[0].rot, out[0].tilt, out[0].psi = 12.5, 77.25, -33.0
    out[0].xoff, out[0].yoff, out[0].iclass = 3.0, -2.0, 4
    out[0].dLL, out[0].pmax, out[0].nr_significant_coarse, out[0].norm_correction = -123456.75, 0.625, 17, 1.03125
    out[1].psi, out[1].xoff, out[1].iclass, out[1].pmax = 90.0, 0.5, 1, 1e-5
    T.check(lib.t2x_star_store_results(h, 1, 2, out, 3))
    q = tmp_path / "run_it001_data.star"
    T.check(lib.t2x_star_write(h, str(q).encode()))
    lib.t2x_star_free(h)
    text = q.read_text()
    assert "data_optics" in text and "data_particles" in text and "_rlnClassNumber #" in text
    assert '"mic 2.mrc"' in text and '"group two"' in text
    h = read_star(lib, q)

    def get(label, row, block=b"particles"):
        v = C.c_double()
        T.check(lib.t2x_star_get_double(h, block, label, row, C.byref(v)), label.decode())
        return v.value

    assert (get(b"rlnAngleRot", 1), get(b"rlnAngleTilt", 1), get(b"rlnAnglePsi", 1)) == (12.5, 77.25, -33.0)
    assert abs(get(b"rlnOriginXAngst", 1) - 3.0 * 0.88) < 1e-9 and abs(get(b"rlnOriginYAngst", 1) + 1.76) < 1e-9
    assert get(b"rlnClassNumber", 1) == 4 and get(b"rlnNrOfSignificantSamples", 1) == 17
    assert get(b"rlnLogLikeliContribution", 1) == -123456.75 and get(b"rlnMaxValueProbDistribution", 1) == 0.625
    assert get(b"rlnNormCorrection", 1) == 1.03125
    assert abs(get(b"rlnOriginXAngst", 2) - 0.55) < 1e-9 and get(b"rlnMaxValueProbDistribution", 2) == 1e-5
    # untouched row and columns survive; columns new to the file exist in every row
    assert get(b"rlnAnglePsi", 0) == 30.0 and get(b"rlnDefocusU", 0) == 12000.5
    assert get(b"rlnClassNumber", 0) == 0 and get(b"rlnImagePixelSize", 1, b"optics") == 0.88
    lib.t2x_star_free(h)


def test_star_list_block_and_escaping(tmp_path):
    lib = T.load()
    p = tmp_path / "job.star"
    p.write_text("# version 30001\n\ndata_job\n\n_rlnJobTypeLabel             relion.refine3d\n"
                 "_rlnJobIsContinue                       0\n \n\n# version 30001\n\ndata_joboptions_values\n\n"
                 "loop_ \n_rlnJobOptionVariable #1 \n_rlnJobOptionValue #2 \n"
                 'auto_local_sampling "1.8 degrees" \n   fn_cont         "" \nnr_threads         8 \n \n')
    h = read_star(lib, p)
    s = C.create_string_buffer(64)
    T.check(lib.t2x_star_get_string(h, b"job", b"rlnJobTypeLabel", 0, s, 64))
    assert s.value == b"relion.refine3d"
    v = C.c_double()
    T.check(lib.t2x_star_get_double(h, b"job", b"rlnJobIsContinue", 0, C.byref(v)))
    assert v.value == 0
    rows = [(b"auto_local_sampling", b"1.8 degrees"), (b"fn_cont", b""), (b"nr_threads", b"8")]
    for i, (k, val) in enumerate(rows):
        T.check(lib.t2x_star_get_string(h, b"joboptions_values", b"rlnJobOptionVariable", i, s, 64))
        assert s.value == k
        T.check(lib.t2x_star_get_string(h, b"joboptions_values", b"rlnJobOptionValue", i, s, 64))
        assert s.value == val
    q = tmp_path / "job2.star"
    T.check(lib.t2x_star_write(h, str(q).encode()))
    lib.t2x_star_free(h)
    h = read_star(lib, q)
    for i, (k, val) in enumerate(rows):
        T.check(lib.t2x_star_get_string(h, b"joboptions_values", b"rlnJobOptionValue", i, s, 64))
        assert s.value == val
    lib.t2x_star_free(h)


def test_star_errors(tmp_path):
    lib = T.load()
    p = tmp_path / "bad.star"
    p.write_text("data_x\n\nloop_\n_rlnA #1\n_rlnB #2\n_rlnC #3\n1 2 3 4\n")
    h = C.c_void_p()
    assert lib.t2x_star_read(str(p).encode(), C.byref(h)) == T.T2X_EINVAL
    assert b"more columns" in lib.t2x_last_error()
    assert lib.t2x_star_read(str(tmp_path / "missing.star").encode(), C.byref(h)) == T.T2X_EINVAL

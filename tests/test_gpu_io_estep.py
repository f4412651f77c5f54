"""The E-step driven from files: particles written as an MRC stack and a RELION 3.1 particles.star,
read back through include/t2x_io.h, give exactly the E-step inputs and results of the in-memory
path, and the results land in the run_itNNN_data.star columns."""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def write_star(path, wl, stack_name):
    px = wl.cfg.pixel_size
    c0 = wl.pin[0].ctf
    lines = ["", "# version 30001", "", "data_optics", "", "loop_",
             "_rlnOpticsGroup #1", "_rlnVoltage #2", "_rlnSphericalAberration #3", "_rlnAmplitudeContrast #4",
             "_rlnImagePixelSize #5",
             f"1 {c0.voltage_kv!r} {c0.cs_mm!r} {c0.q0!r} {px!r}", "", "# version 30001", "", "data_particles", "",
             "loop_"]
    cols = ["rlnImageName", "rlnOpticsGroup", "rlnDefocusU", "rlnDefocusV", "rlnDefocusAngle", "rlnCtfBfactor",
            "rlnCtfScalefactor", "rlnPhaseShift", "rlnOriginXAngst", "rlnOriginYAngst", "rlnAngleRot",
            "rlnAngleTilt", "rlnAnglePsi", "rlnNormCorrection"]
    lines += [f"_{c} #{i + 1}" for i, c in enumerate(cols)]
    for i in range(wl.n):
        p = wl.pin[i]
        assert (p.ctf.voltage_kv, p.ctf.cs_mm, p.ctf.q0) == (c0.voltage_kv, c0.cs_mm, c0.q0)
        vals = [f"{i + 1:06d}@{stack_name}", "1", repr(p.ctf.defocus_u), repr(p.ctf.defocus_v),
                repr(p.ctf.defocus_angle), repr(p.ctf.bfactor), repr(p.ctf.scale), repr(p.ctf.phase_shift),
                repr(p.old_offset_x * px), repr(p.old_offset_y * px), repr(p.rot), repr(p.tilt), repr(p.psi),
                repr(p.norm_correction)]
        lines.append(" ".join(vals))
    path.write_text("\n".join(lines) + "\n \n")


def test_estep_from_star_and_mrc(tmp_path, gpu_device):
    lib = T.load()
    wl = H.Workload("refine64", n=16, seed_offset=11, precision=T.T2X_PREC_FP16X3)
    N, n = wl.N, wl.n
    stack = tmp_path / "stack.mrcs"
    T.check(lib.t2x_mrc_write(str(stack).encode(), N, N, n, wl.cfg.pixel_size, T.fptr(wl.images)))
    star = tmp_path / "particles.star"
    write_star(star, wl, "stack.mrcs")
    ref = wl.run_gpu(dev=gpu_device)

    h = C.c_void_p()
    T.check(lib.t2x_star_read(str(star).encode(), C.byref(h)))
    pin2 = (T.ParticleIn * n)()
    T.check(lib.t2x_star_particles(h, 0, n, pin2, None))
    images2 = np.zeros((n, N, N), np.float32)
    idx = C.c_long()
    name = C.create_string_buffer(512)
    for i in range(n):
        T.check(lib.t2x_star_image_ref(h, i, C.byref(idx), name, 512))
        T.check(lib.t2x_mrc_read_sections(str(tmp_path / name.value.decode()).encode(), idx.value - 1, 1,
                                          T.fptr(images2[i])))
    np.testing.assert_array_equal(images2, wl.images)
    for i in range(n):
        for f, _ in T.CtfParams._fields_:
            assert getattr(pin2[i].ctf, f) == getattr(wl.pin[i].ctf, f), (i, f)
        for f in ("rot", "tilt", "psi", "norm_correction"):
            assert getattr(pin2[i], f) == getattr(wl.pin[i], f), (i, f)
        assert abs(pin2[i].old_offset_x - wl.pin[i].old_offset_x) < 1e-12
        assert abs(pin2[i].old_offset_y - wl.pin[i].old_offset_y) < 1e-12

    wl.images, wl.pin = images2, pin2
    got = wl.run_gpu(dev=gpu_device)
    assert np.array_equal(got["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(got["out"]["dLL"], ref["out"]["dLL"])

    outs = (T.ParticleOut * n)()
    for i in range(n):
        for f in ("rot", "tilt", "psi", "xoff", "yoff", "iclass", "pmax", "dLL", "norm_correction",
                  "nr_significant_coarse"):
            setattr(outs[i], f, got["out"][f][i].item())
    T.check(lib.t2x_star_store_results(h, 0, n, outs, 3))
    out_star = tmp_path / "run_it001_data.star"
    T.check(lib.t2x_star_write(h, str(out_star).encode()))
    lib.t2x_star_free(h)
    T.check(lib.t2x_star_read(str(out_star).encode(), C.byref(h)))
    v = C.c_double()
    for i in range(n):
        T.check(lib.t2x_star_get_double(h, b"particles", b"rlnClassNumber", i, C.byref(v)))
        assert v.value == got["out"]["iclass"][i]
        T.check(lib.t2x_star_get_double(h, b"particles", b"rlnAnglePsi", i, C.byref(v)))
        assert abs(v.value - got["out"]["psi"][i]) < 1e-6
        T.check(lib.t2x_star_get_double(h, b"particles", b"rlnOriginXAngst", i, C.byref(v)))
        assert abs(v.value - got["out"]["xoff"][i] * wl.cfg.pixel_size) < 1e-6
    lib.t2x_star_free(h)

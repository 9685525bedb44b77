"""The C-ABI boundary on the CPU: the library loads, exports every function
include/kp_image.h declares, its struct layouts match the ctypes mirror (as a
C compiler lays them out), the default initialisers and the option layer work
without a device, and the option layer agrees with the oracle's restatement of
the reference's host rules.  No compute call is made here."""
from __future__ import annotations

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from kpolaris_amd import abi, batch_jobs, image
from oracle import host_rules as hr

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "kp_image.h"
PARAMS = ROOT / "kpolaris_amd" / "params"

STRUCTS = {
    "kp_camera_params": abi.CameraParams,
    "kp_emission_selection": abi.EmissionSelection,
    "kp_pass_a_params": abi.PassAParams,
    "kp_riaf_model": abi.RIAFModel,
    "kp_image_data": abi.ImageData,
    "kp_run_options": abi.RunOptions,
    "kp_image_stats": abi.ImageStats,
    "kp_batch_job": abi.BatchJob,
    "kp_slow_light": abi.SlowLight,
    "kp_slow_light_stream": abi.SlowLightStream,
    "kp_slow_light_schedule": abi.SlowLightSchedule,
    "kp_slow_light_job": abi.SlowLightJob,
}


def test_library_exports_every_header_function():
    L = image.library()
    names = image.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.kp_abi_version() == abi.ABI_VERSION


def test_struct_layouts_match_a_c_compiler(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_defaults_and_options_without_a_device():
    L = image.library()
    p = abi.PassAParams()
    L.kp_pass_a_params_default(C.byref(p))
    m = abi.RIAFModel()
    L.kp_riaf_model_default(C.byref(m))
    o = abi.RunOptions()
    L.kp_run_options_default(C.byref(o))
    assert o.split_transport == 1 and o.persistent == 1 and o.precision == abi.PRECISION_FP64
    assert o.fp64_radius == abi.FP64_RADIUS_DEFAULT
    assert L.kp_device_count() >= 0
    opts = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    with pytest.raises(image.KPError):
        opts.set("nx", "not-a-number")


@pytest.mark.parametrize("par", ["riaf_m87_c1.par", "riaf_m87_c2.par"])
def test_option_layer_equals_host_rules(par):
    """The C++ option layer against the oracle's restatement of
    tools/kpolaris_model_image_options.cpp and tools/kpolaris_model_image.cpp."""
    p, m, scale, sign = image.Options(parameter_file=str(PARAMS / par)).resolve_riaf()
    opt = hr.parse_parameter_text((PARAMS / par).read_text())
    rp, rm, rscale, rsign = hr.resolve_riaf(opt)
    for name, _ in abi.PassAParams._fields_:
        if name in ("camera", "emission_selection"):
            continue
        assert getattr(p, name) == getattr(rp, name), name
    for name, _ in abi.CameraParams._fields_:
        assert getattr(p.camera, name) == getattr(rp.camera, name), name
    for name, _ in abi.RIAFModel._fields_:
        assert getattr(m, name) == getattr(rm, name), name
    assert scale == rscale and sign == rsign


def test_batch_job_draws_are_seeded_and_in_range():
    a = batch_jobs.draw(seed=5, n=64)
    b = batch_jobs.draw(seed=5, n=64)
    assert a == b
    for j in a:
        assert j["nx"] in batch_jobs.SIZES and j["nx"] == j["ny"]
        assert j["spin"] in batch_jobs.SPINS and j["freq"] in batch_jobs.FREQS
        assert 10.0 <= j["inclination_deg"] <= 90.0 and 0.3 <= j["ne_scale"] <= 3.0
    p, m = batch_jobs.resolve(a[0])
    assert p.camera.nx == a[0]["nx"] and m.freq_cgs == a[0]["freq"]
    rh = 1.0 + np.sqrt(1.0 - a[0]["spin"] ** 2)
    assert p.inner_radius == pytest.approx(1.05 * rh)


def test_plain_c_caller_compiles_and_links():
    """A C99 program that includes only include/kp_image.h builds and links
    against the library (the ABI is plain C)."""
    from tests.c_caller import build as cb
    exe = cb.build()
    assert exe.exists()

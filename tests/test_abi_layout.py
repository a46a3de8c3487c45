"""The boundary records keep the reference's x86-64 layout.

Facts come from the reference's own ABI contract file
(tests/ffi/cuda_abi_facts.inc), extracted to tests/golden/abi_facts.json.
Both the C header (compiled with gcc) and the ctypes mirror are checked.
"""
import ctypes
import json
import os
import subprocess
import tempfile

import pytest

from fj import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FACTS = json.load(open(os.path.join(HERE, "golden", "abi_facts.json")))


def test_ctypes_mirror_matches_reference_facts():
    checked = 0
    for name, (size, align) in FACTS["types"].items():
        if name.endswith("*") or name == "FjAbortQuery":
            continue
        t = abi.ALL_TYPES[name]
        assert ctypes.sizeof(t) == size, name
        assert ctypes.alignment(t) == align, name
        checked += 1
    for name, field, off in FACTS["fields"]:
        assert getattr(abi.ALL_TYPES[name], field).offset == off, (name, field)
    assert checked == 43 and len(FACTS["fields"]) == 414


def test_c_header_matches_reference_facts():
    lines = ['#include <stddef.h>', '#include "fj/juicer_cuda_api.h"']
    for name, (size, align) in FACTS["types"].items():
        lines.append(f'_Static_assert(sizeof({name}) == {size} && _Alignof({name}) == {align}, "{name}");')
    for name, field, off in FACTS["fields"]:
        lines.append(f'_Static_assert(offsetof({name}, {field}) == {off}, "{name}.{field}");')
    for name, value in FACTS["values"].items():
        lines.append(f'_Static_assert({name} == {value}, "{name}");')
    lines.append("int main(void) { return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "abi.c")
        open(src, "w").write("\n".join(lines))
        r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-c", src, "-o",
                            os.path.join(d, "abi.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]


def test_header_compiles_as_cpp():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "abi.cpp")
        open(src, "w").write('#include "fj/juicer_cuda_api.h"\nint main() { FjPreparedHostData p{}; return (int)p.route; }\n')
        r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-c", src, "-o",
                            os.path.join(d, "abi.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]


# Call scope and frame values (tests/ffi/prepared_call_abi_facts.inc, extracted
# to tests/golden/call_abi_facts.json). Its two values name the reference's
# test-only API (juicer_test_api.h) and are not part of the boundary.
CALL_FACTS = json.load(open(os.path.join(HERE, "golden", "call_abi_facts.json")))


def test_call_scope_ctypes_mirror_matches_reference_facts():
    for name, (size, align) in CALL_FACTS["types"].items():
        t = abi.ALL_TYPES[name]
        assert ctypes.sizeof(t) == size, name
        assert ctypes.alignment(t) == align, name
    for name, field, off in CALL_FACTS["fields"]:
        assert getattr(abi.ALL_TYPES[name], field).offset == off, (name, field)
    assert len(CALL_FACTS["types"]) == 8 and len(CALL_FACTS["fields"]) == 50


def test_call_scope_header_matches_reference_facts():
    lines = ['#include <stddef.h>', '#include "fj/juicer_cuda_call_scope.h"', '#include "fj/fj_mi355x_ext.h"']
    for name, (size, align) in CALL_FACTS["types"].items():
        lines.append(f'_Static_assert(sizeof({name}) == {size} && _Alignof({name}) == {align}, "{name}");')
    for name, field, off in CALL_FACTS["fields"]:
        lines.append(f'_Static_assert(offsetof({name}, {field}) == {off}, "{name}.{field}");')
    # The frame-math records' sizes follow from the offsets above.
    lines.append('_Static_assert(sizeof(FjDirFrameMath) == 512, "FjDirFrameMath");')
    lines.append('_Static_assert(sizeof(FjGrainFrameMath) == 552, "FjGrainFrameMath");')
    lines.append('_Static_assert(sizeof(FjEffectsFrameMath) == 208, "FjEffectsFrameMath");')
    lines.append("int main(void) { return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "abi.c")
        open(src, "w").write("\n".join(lines))
        r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", src,
                            "-o", os.path.join(d, "abi.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
    for name in ("FjDirFrameMath", "FjGrainFrameMath", "FjEffectsFrameMath"):
        assert ctypes.sizeof(abi.ALL_TYPES[name]) == {"FjDirFrameMath": 512, "FjGrainFrameMath": 552,
                                                      "FjEffectsFrameMath": 208}[name]


# Import records of the optional host functions (the reference's C11 layout
# assertions in tests/ffi/prepared-values.c, extracted to
# tests/golden/boundary_abi_facts.json).
BOUNDARY_FACTS = json.load(open(os.path.join(HERE, "golden", "boundary_abi_facts.json")))


def test_boundary_import_records_match_reference_facts():
    for name, (size, align) in BOUNDARY_FACTS["types"].items():
        t = abi.ALL_TYPES[name]
        assert ctypes.sizeof(t) == size and ctypes.alignment(t) == align, name
    for name, field, off in BOUNDARY_FACTS["fields"]:
        assert getattr(abi.ALL_TYPES[name], field).offset == off, (name, field)
    assert len(BOUNDARY_FACTS["types"]) == 6 and len(BOUNDARY_FACTS["fields"]) == 36
    lines = ['#include <stddef.h>', '#include "fj/juicer_diffusion_api.h"', '#include "fj/juicer_dir_boundary_api.h"']
    for name, (size, align) in BOUNDARY_FACTS["types"].items():
        lines.append(f'_Static_assert(sizeof({name}) == {size} && _Alignof({name}) == {align}, "{name}");')
    for name, field, off in BOUNDARY_FACTS["fields"]:
        lines.append(f'_Static_assert(offsetof({name}, {field}) == {off}, "{name}.{field}");')
    lines.append("int main(void) { return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "abi.c")
        open(src, "w").write("\n".join(lines))
        r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", src,
                            "-o", os.path.join(d, "abi.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]


def test_tc_lut_input_layout():
    """FjFilmTcLutInput has the reference's x86-64 layout
    (tests/ffi/juicer_test_values.h:964-973: three 16-byte spans, a float[3]
    white, the blur, two tags, a span, two flags, a float[2] centre, a span)."""
    import ctypes as C
    from fj import abi
    t = abi.FjFilmTcLutInput
    want = {"spectra": 0, "sensitivity_rgb": 16, "reference_illuminant": 32, "projection_white_xyz": 48,
            "spectral_blur": 60, "method": 64, "apply_surface": 68, "surface_rgb": 72, "compression_active": 88,
            "hull_available": 92, "hull_center_xy": 96, "hull_xy": 104}
    for name, off in want.items():
        assert getattr(t, name).offset == off, name
    assert C.sizeof(t) == 120


def test_print_preparation_layouts():
    """The print-preparation records have the reference's layouts (the
    compile-time assertions of rust film-juicer-plugin/src/print_bridge.rs:50-54,
    644-679), both in ctypes and in the C header compiled by gcc."""
    import ctypes as C
    import subprocess
    import tempfile
    from fj import abi
    want = {
        "FjPrintIlluminantParams": (44, {"edges_nm": 0, "transitions_nm": 16, "cmy_cc": 32}),
        "FjPrintIlluminantFailure": (8, {"reason": 0, "sample": 4}),
        "FjIlluminant": (324, {"samples": 0}),
        "FjPrintPreflashInput": (48, {"film_base_density": 0, "filtered_illuminant": 16,
                                      "print_sensitivity_cmy": 32}),
        "FjPrintPreflashRaw": (12, {"cmy": 0}),
        "FjPrintBalanceInput": (104, {"film_base_density": 0, "film_channel_density_cmy": 16,
                                      "print_sensitivity_rgb": 32, "filtered_main_illuminant": 48,
                                      "baseline_density_cmy": 64, "compensated_density_cmy": 76,
                                      "reference_hash": 88, "normalization_mode": 96}),
        "FjPrintBalance": (56, {"baseline_raw_rgb": 0, "compensated_raw_rgb": 12, "factor_midgray": 24,
                                "factor_midgray_comp": 28, "normalizer": 32,
                                "filtered_main_illuminant_hash": 40, "hash": 48}),
        "FjPrintBalanceFailure": (8, {"reason": 0, "channel": 4}),
    }
    lines = ["#include <stddef.h>", "#include \"fj/fj_mi355x_ext.h\""]
    for name, (size, fields) in want.items():
        t = getattr(abi, name)
        assert C.sizeof(t) == size, name
        lines.append(f"_Static_assert(sizeof({name}) == {size}, \"{name}\");")
        for field, off in fields.items():
            assert getattr(t, field).offset == off, (name, field)
            lines.append(f"_Static_assert(offsetof({name}, {field}) == {off}, \"{name}.{field}\");")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "print.c")
        open(src, "w").write("\n".join(lines) + "\n")
        r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", src,
                            "-o", os.path.join(d, "print.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]

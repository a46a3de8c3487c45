"""GPU parity of the per-pixel chain's method branches the prepared cases do
not take, against the FP64 oracle, through fj_cuda_render on cuda:0:

* input CCTF decode (JuicerCudaDeviceHelpers.cuh:179-225) for the BT.2020 and
  sRGB/Rec.709 input spaces, on both routes;
* every output encoding the reference defines (its nine output spaces: CCTF
  kind, parameters and DWG-to-output matrix, the words of
  tests/scanner/output_encoding_native.inc), with and without the
  input-is-output-space flag;
* Arctic reconstruction (FJ_RAW_ARCTIC_2026_BETA04, recipe.rs:1274-1278) on a
  TC LUT that fjx_build_tc_lut builds with the Arctic neutral response.

Gates as the rest of the GPU parity suite: linear output max |err| <= 2e-4,
encoded output |err| <= 2e-4 + 3e-4 |ref| per sample
(tests/ofx/render_assertions.cpp:16), alpha bit-exact.
"""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import tc_lut_cases as T  # noqa: E402
from fj import abi, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402
from test_gpu_pixel_parity import check, render_gpu  # noqa: E402

PIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pixel")


def f32(words):
    return np.asarray(words, dtype=np.uint32).view(np.float32)


@pytest.mark.parametrize("space", [abi.FJ_INPUT_BT2020, abi.FJ_INPUT_SRGB_REC709], ids=["bt2020", "srgb"])
@pytest.mark.parametrize("case", ["NegativeDirectHanatos", "PrintPreflash"])
def test_input_cctf_decode(case, space):
    p = prepared.simple_view(prepared.load(case), encode=False)
    fe = p.record.film_exposure
    fe.flags |= abi.FJ_FILM_DECODE_CCTF
    fe.input_color_space = space
    # Encoded-domain inputs: mostly [0, 1], a few above 1 and below 0.
    src = synth.frame(96, 48, seed=21)
    rgb = src[..., :3]
    src[..., :3] = np.clip(rgb / (1.0 + rgb), -0.05, 1.2)
    src[0, :4, :3] = [[-0.5, 0.0, 0.02], [0.04045, 0.0812428791, 1.0], [1.5, 2.0, 0.5], [1e-7, 0.3, 0.9]]
    got, out = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, None, lin, src, True)
    print(f"{case} decode space {space}: max {err.max():.3e}")
    assert out.flags == 0


@pytest.mark.parametrize("io", [False, True], ids=["native", "input-is-output"])
@pytest.mark.parametrize("entry", range(9))
def test_output_encodings(entry, io):
    d = np.load(os.path.join(PIX, "encoding_entries.npz"))
    p = prepared.simple_view(prepared.load("NegativePrintHanatos"), encode=True)
    oc = p.record.output_color
    if io:
        oc.flags |= abi.FJ_COLOR_INPUT_IS_OUTPUT_SPACE
    e = oc.encoding
    e.schema = abi.FJ_OUTPUT_ENCODING_SCHEMA
    e.cctf_kind = int(d["kind"][entry])
    e.gamma, e.a, e.b, e.c, e.d, e.linear_cutoff = f32(d["params"][entry]).tolist()
    e.dwg_to_output[:] = f32(d["dwg_to_output"][entry]).tolist()
    src = synth.frame(80, 40, seed=23)
    got, out = render_gpu(p, src)
    dst, _, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, dst, None, src, False)
    print(f"encoding entry {entry} kind {e.cctf_kind} io {io}: max {err.max():.3e}")
    assert out.flags == 0


def _arctic_table():
    """An Arctic TC LUT built on the GPU from a procedural case of the
    reference's TC-LUT fixture (method 2, the Arctic neutral response)."""
    from test_gpu_tc_lut import gpu_build

    cases = [c for c in T.procedural_cases() if c["built"] and T.inputs(c)["method"] == 2]
    assert cases, "the fixture holds built Arctic cases"
    case = cases[0]
    st, text, table = gpu_build(**T.inputs(case))
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    assert oracle.fnv1a64(table) == case["digest"]
    return table


@pytest.mark.parametrize("case", ["NegativeDirectHanatos", "NegativePrintHanatos"])
def test_arctic_reconstruction(case):
    table = _arctic_table()
    p = prepared.simple_view(prepared.load(case), encode=False)
    fe = p.record.film_exposure
    assert fe.method != abi.FJ_RAW_MALLETT_2019
    fe.method = abi.FJ_RAW_ARCTIC_2026_BETA04
    fe.tc_lut_rgba = p.fspan(table.ravel())
    fe.tc_lut_hash = int(oracle.fnv1a64(table))
    src = synth.frame(120, 64, seed=25)
    got, out = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, None, lin, src, True)
    print(f"{case} Arctic: max {err.max():.3e} mean {err.mean():.3e}")
    assert out.flags == 0

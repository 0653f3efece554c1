"""Reverse pdfs of the oracle's materials (SurfaceShader.Pdf(..).PdfReverse,
DiffuseMaterial.cs:63-71,95-103,167-188, GenericMaterial.cs:152-338), which only
the bidirectional MIS weights read.

The reverse pdf of a pair (out, in) is the density of sampling `out` when
`in` is the outgoing direction. Per lobe that is the forward density of the
swapped pair, so:
  * DiffuseMaterial (opaque and transmitter): reverse(out, in) == forward(in, out);
  * GenericMaterial with Metallic 1 (only the microfacet reflection lobe, its
    selection probability 1): reverse(out, in) == forward(in, out);
  * GenericMaterial with SpecularTransmittance 1 and Metallic 0 for a pair on
    opposite sides whose reflection half vector points up (so the reflection
    lobe's density of `in` is zero both ways): only the transmission lobe is
    left, and the reference weights the reverse lobe with the selection
    probability of the forward direction (selectTransmitIn repeats
    (1 - fresnelT) of `out`, GenericMaterial.cs:311), so
    reverse(out, in) / forward(in, out) == (1 - F(out . h)) / (1 - F(in . h))
    with F the dielectric Fresnel term and h the transmission half vector.
A transmission lobe with the forward instead of the swapped IOR ratio, a
reflection lobe evaluated for the wrong direction or another selection weight
fails one of these. BSDF samples carry the reverse pdf of their direction."""
import ctypes as C

import numpy as np

import oracle_lib as ol
from test_oracle import NORMAL, UV, arr, diffuse_desc, f3, generic_desc, sample

L = ol.lib()


def pdf_pair(m, out_dir, in_dir):
    fwd, rev = C.c_float(), C.c_float()
    L.or_material_pdf_pair(C.byref(m), None, f3(*NORMAL), f3(*UV), f3(*out_dir), f3(*in_dir), C.byref(fwd),
                           C.byref(rev))
    return fwd.value, rev.value


def directions(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_diffuse_reverse_is_swapped_forward():
    outs, ins = directions(300, 1), directions(300, 2)
    for transmitter in (False, True):
        m = diffuse_desc(transmitter=transmitter)
        nonzero = 0
        for o, i in zip(outs, ins):
            fwd_oi, rev_oi = pdf_pair(m, o, i)
            fwd_io, rev_io = pdf_pair(m, i, o)
            assert rev_oi == fwd_io and rev_io == fwd_oi
            nonzero += rev_oi > 0
        assert nonzero > 100


def test_metallic_reverse_is_swapped_forward():
    outs, ins = directions(400, 3), directions(400, 4)
    for rough in (0.1, 0.4, 0.9):
        m = generic_desc(rough=rough, metallic=1.0)
        nonzero = 0
        for o, i in zip(outs, ins):
            fwd_io, _ = pdf_pair(m, i, o)
            _, rev_oi = pdf_pair(m, o, i)
            assert rev_oi == fwd_io
            nonzero += rev_oi > 0
        assert nonzero > 100


def test_transmission_reverse_pdf():
    ior = 1.5
    m = generic_desc(rough=0.5, metallic=0.0, trans=1.0, ior=ior)
    up = np.array(NORMAL, np.float64)
    outs, ins = directions(200000, 5), directions(200000, 6)
    # out above, in below, reflection half vector up
    keep = (outs @ up > 0.05) & (ins @ up < -0.05) & ((outs + ins) @ up > 0.05)
    checked = 0
    for o, i in zip(outs[keep], ins[keep]):
        o64, i64 = o.astype(np.float64), i.astype(np.float64)
        fwd_io, _ = pdf_pair(m, i, o)
        _, rev_oi = pdf_pair(m, o, i)
        h = o64 + i64 * ior  # the transmission half vector (out above: eta = ior)
        h /= np.linalg.norm(h)
        h = h if h @ up > 0 else -h
        f_out = L.or_fresnel_dielectric(float(o64 @ h), 1.0, ior)
        f_in = L.or_fresnel_dielectric(float(i64 @ h), 1.0, ior)
        if fwd_io < 1e-3 or f_in > 0.9:
            continue  # (the fp32 half vector's rounding moves F visibly near grazing)
        np.testing.assert_allclose(rev_oi / fwd_io, (1.0 - f_out) / (1.0 - f_in), rtol=2e-4)
        checked += 1
        if checked == 300:
            break
    assert checked == 300


def test_samples_carry_reverse_pdf():
    rng = np.random.default_rng(6)
    for m in (diffuse_desc(), diffuse_desc(transmitter=True), generic_desc(rough=0.3, metallic=0.5),
              generic_desc(rough=0.3, trans=0.8)):
        o = directions(1, 7)[0]
        o[1] = abs(o[1])
        for _ in range(50):
            pc, p2 = float(rng.random()), rng.random(2).astype(np.float32)
            d, pdf, _ = sample(m, o, pc, p2)
            if pdf == 0.0:
                continue
            out = arr(8)
            L.or_material_sample(C.byref(m), None, f3(*NORMAL), f3(*UV), f3(*o), pc, f3(*p2), out)
            fwd, rev = pdf_pair(m, o, tuple(d))
            assert out[7] == rev and out[3] == fwd

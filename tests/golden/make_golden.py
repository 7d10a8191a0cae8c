#!/usr/bin/env python3
"""Writes tests/golden/golden.npz: inputs and expected outputs computed with
numpy / scipy (float64), independent of the oracle.

Cases:
- lines: C2C / R2C lengths 1..13 and the config line lengths, batches 1 and 3
  (the reference's dft_oracle cases, crates/pencil-fft/src/lib.rs:900-914);
- r2r: the eight FFTW kinds via scipy.fft.dct/dst(type, norm=None), which
  uses the FFTW REDFT/RODFT unnormalized convention, plus the DHT
  (Re F - Im F);
- C2R-oriented routes (C2rPlan, crates/pencil-fft/src/distributed/c2r.rs:72-361):
  numpy.fft.rfftn spectra of fixture fields, raw C2R output = field x the
  selected extents, in the reversed memory order;
- routes: the reference's fixture case list
  (tools/fftw-reference/README.md:271-294) with its input formulas
  (tools/fftw-reference/generate.jl:201-250), transformed with numpy.fft /
  scipy.fft over the selected axes and laid out in the plan's output memory
  order (reversed axes, build_route in distributed.rs:2306-2373).
Run from the repository root: python tests/golden/make_golden.py
"""
import itertools
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def complex_value(coords, seed):
    re = 0.173 + 0.019 * seed
    im = -0.271 - 0.013 * seed
    for axis, c in enumerate(coords, start=1):
        re += (0.071 + 0.009 * axis) * (c + 1)
        im += (0.053 + 0.011 * axis) * (c + 2)
        re += 0.017 * np.sin((axis + 2) * (c + 1) + seed * 0.07)
        im += 0.023 * np.cos((axis + 3) * (c + 2) - seed * 0.05)
    for left in range(1, len(coords)):
        for right in range(left + 1, len(coords) + 1):
            re += 0.0067 * left * (right + 1) * (coords[left - 1] + 1) * (coords[right - 1] + 2)
            im -= 0.0049 * (left + 1) * right * (coords[left - 1] + 2) * (coords[right - 1] + 1)
    return re + 1j * im


def real_value(coords, seed):
    v = 0.417 + 0.023 * seed
    for axis, c in enumerate(coords, start=1):
        v += (0.087 + 0.013 * axis) * (c + 1)
        v += 0.031 * np.sin((axis + 1) * (c + 2) + seed * 0.09)
    for left in range(1, len(coords)):
        for right in range(left + 1, len(coords) + 1):
            v += 0.0083 * (left + 1) * (right + 2) * (coords[left - 1] + 1) * (coords[right - 1] + 2)
    return v


def field(shape, extra, seed, complex_):
    out = np.empty((extra,) + tuple(shape), dtype=np.complex128 if complex_ else np.float64)
    for e in range(extra):
        for idx in itertools.product(*[range(n) for n in shape]):
            out[(e,) + idx] = (complex_value if complex_ else real_value)(list(idx), seed + e)
    return out


def out_order(x):
    """[extra][axis0..axisN-1] -> [extra][axisN-1..axis0] (output pencil)."""
    nd = x.ndim - 1
    return np.ascontiguousarray(np.transpose(x, (0,) + tuple(range(nd, 0, -1))))


R2R = {1: (sf.dct, 1), 2: (sf.dct, 2), 3: (sf.dct, 3), 4: (sf.dct, 4),
       5: (sf.dst, 1), 6: (sf.dst, 2), 7: (sf.dst, 3), 8: (sf.dst, 4)}


def r2r_axis(x, kind, axis):
    if kind == 9:
        f = np.fft.fft(x, axis=axis)
        return f.real - f.imag
    fn, t = R2R[kind]
    return fn(x, type=t, axis=axis, norm=None)


def main():
    g = {}
    rng = np.random.default_rng(20261202)
    for n in [1, 2, 3, 4, 5, 7, 8, 9, 11, 12, 13, 16, 48, 60, 64, 84, 96, 112, 128, 160, 240, 256, 384, 385, 512]:
        for b in (1, 3):
            x = rng.uniform(-1, 1, (b, n)) + 1j * rng.uniform(-1, 1, (b, n))
            g[f"c2c_n{n}_b{b}_in"] = x
            g[f"c2c_n{n}_b{b}_fwd"] = np.fft.fft(x, axis=-1)
            g[f"c2c_n{n}_b{b}_inv"] = np.fft.ifft(x, axis=-1)
            xr = rng.uniform(-1, 1, (b, n))
            g[f"r2c_n{n}_b{b}_in"] = xr
            g[f"r2c_n{n}_b{b}_fwd"] = np.fft.rfft(xr, axis=-1)
    for kind in range(1, 10):
        for n in [2, 3, 4, 5, 6, 7, 8, 12, 13, 16, 60, 96, 240, 256, 384]:
            x = rng.uniform(-1, 1, (3, n))
            g[f"r2r_k{kind}_n{n}_in"] = x
            g[f"r2r_k{kind}_n{n}_fwd"] = r2r_axis(x, kind, -1)
    # routes: (name, kind, shape, extra, selection or kinds)
    c2c = [((3, 4), 1, None), ((3, 2, 5), 6, None), ((2, 1, 3, 4), 2, None), ((2, 3, 2, 3), 1, []),
           ((2, 3, 2, 3), 1, [0, 3])]
    for i, (shape, extra, sel) in enumerate(c2c):
        x = field(shape, extra, 1.0, True)
        axes = tuple(a + 1 for a in (range(len(shape)) if sel is None else sel))
        y = np.fft.fftn(x, axes=axes) if axes else x.copy()
        g[f"route_c2c{i}_in"] = x
        g[f"route_c2c{i}_fwd"] = out_order(y)
        g[f"route_c2c{i}_meta"] = np.array(list(shape) + [extra] + [-1] + (list(sel) if sel is not None else
                                                                          list(range(len(shape)))))
    r2c = [((3, 1), 1, None), ((3, 2), 1, None), ((3, 1, 4), 1, None), ((3, 1, 5), 6, None),
           ((2, 1, 3, 3), 2, None), ((2, 3, 2, 3), 1, [0]), ((2, 3, 2, 3), 1, [0, 2]), ((2, 3, 2, 3), 1, [0, 3]),
           ((2, 3, 4), 2, [0, 2])]
    for i, (shape, extra, sel) in enumerate(r2c):
        x = field(shape, extra, 2.0, False)
        axes = tuple(a + 1 for a in (range(len(shape)) if sel is None else sel))
        y = np.fft.rfftn(x, axes=axes)  # reduces the last selected axis
        g[f"route_r2c{i}_in"] = x
        g[f"route_r2c{i}_fwd"] = out_order(y)
        g[f"route_r2c{i}_meta"] = np.array(list(shape) + [extra] + [-1] + (list(sel) if sel is not None else
                                                                          list(range(len(shape)))))
    # C2R orientation (C2rPlan): reduction axis = first selected axis, the
    # complex input in canonical order, the raw C2R output in the reversed
    # memory order; the input is the spectrum of a fixture field, so the raw
    # output is the field times the product of the selected extents
    c2r = [((3, 4), 1, None), ((3, 2, 5), 2, None), ((4, 3, 5), 1, None), ((2, 3, 2, 3), 1, [1, 3]),
           ((5, 4, 6), 1, [0, 2]), ((6,), 3, None)]
    for i, (shape, extra, sel) in enumerate(c2r):
        x = field(shape, extra, 4.0, False)
        sel_l = list(range(len(shape))) if sel is None else list(sel)
        ra = min(sel_l)
        axes = tuple(a + 1 for a in sel_l if a != ra) + (ra + 1,)
        X = np.fft.rfftn(x, axes=axes)
        g[f"route_c2r{i}_in"] = X
        g[f"route_c2r{i}_fwd"] = out_order(x * float(np.prod([shape[a] for a in sel_l])))
        g[f"route_c2r{i}_meta"] = np.array(list(shape) + [extra] + [-1] + sel_l)
    # heterogeneous plans (distributed/mixed.rs): per-axis codes 0 none,
    # 1..8 R2R, 9 DHT, 10 FFT, 11 RFFT; an R2R axis of complex data transforms
    # the real and imaginary parts separately
    def mixed_axis(y, code, axis):
        if code == 0:
            return y
        if code == 10:
            return np.fft.fft(y, axis=axis)
        if np.iscomplexobj(y):
            return r2r_axis(y.real, code, axis) + 1j * r2r_axis(y.imag, code, axis)
        return r2r_axis(y, code, axis)

    mixed_c2c = [((4, 5, 6), 1, [10, 2, 10]), ((5, 6, 4), 2, [9, 10, 4]), ((3, 4, 5), 1, [0, 10, 5]),
                 ((6, 4, 5), 1, [1, 9, 10]), ((4, 3, 2, 5), 1, [10, 6, 0, 8])]
    for i, (shape, extra, codes) in enumerate(mixed_c2c):
        x = field(shape, extra, 5.0, True)
        y = x.copy()
        for a, c in enumerate(codes):
            y = mixed_axis(y, c, a + 1)
        g[f"route_mixc2c{i}_in"] = x
        g[f"route_mixc2c{i}_fwd"] = out_order(y)
        g[f"route_mixc2c{i}_meta"] = np.array(list(shape) + [extra] + [-1] + list(codes))
    mixed_r2c = [((4, 6, 5), 1, [10, 11, 2]), ((5, 4, 6), 2, [9, 11, 0]), ((6, 5, 8), 1, [2, 10, 11]),
                 ((4, 6, 5, 3), 1, [10, 6, 11, 4])]
    for i, (shape, extra, codes) in enumerate(mixed_r2c):
        x = field(shape, extra, 6.0, False)
        ra = codes.index(11)
        y = x.copy()
        for a in range(len(shape) - 1, ra, -1):  # real prefix, route order
            y = mixed_axis(y, codes[a], a + 1)
        y = np.fft.rfft(y, axis=ra + 1)
        for a in range(ra - 1, -1, -1):
            y = mixed_axis(y, codes[a], a + 1)
        g[f"route_mixr2c{i}_in"] = x
        g[f"route_mixr2c{i}_fwd"] = out_order(y)
        g[f"route_mixr2c{i}_meta"] = np.array(list(shape) + [extra] + [-1] + list(codes))
    r2r = [((3, 4), 1, [9, 9]), ((3, 2, 4), 6, [9, 0, 9]), ((2, 3, 2, 3), 1, [9, 0, 0, 9]),
           ((2, 3, 2, 3), 1, [1, 2, 3, 4]), ((2, 3, 2, 3), 1, [5, 6, 7, 8]), ((6, 5, 4), 2, [2, 5, 4])]
    for i, (shape, extra, kinds) in enumerate(r2r):
        x = field(shape, extra, 3.0, False)
        y = x.copy()
        for a, k in enumerate(kinds):
            if k:
                y = r2r_axis(y, k, a + 1)
        g[f"route_r2r{i}_in"] = x
        g[f"route_r2r{i}_fwd"] = out_order(y)
        g[f"route_r2r{i}_meta"] = np.array(list(shape) + [extra] + [-1] + list(kinds))
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **g)
    print(f"wrote {path}: {len(g)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()

"""Write golden.json: exact results of the reference's test convolutions on
its own integer ramp inputs, computed with Python integers (no floating
point, no FFT, no numpy, no code shared with oracle/).

Inputs follow the reference test programs:
  conv1  tests/hybridconv.cc:49      f_a[j] = (j+a+1) + i((a+1)j+3)
  convC1 tests/hybridconv.cc:49 (-c) the same ramps, centred data
  convh1 tests/hybridconvh.cc:46-50  the same ramps on H modes, Im f[0] = 0
  conv2  tests/hybridconv2.cc:52-57  f_a[i][j] = (i+a+1) + i((a+1)j+3)
  convh2 tests/hybridconvh2.cc:54-63 f_a[i][j] = (I+a+1) + i((a+1)j+3),
         I = i (odd Lx) or i-1 (even Lx), then HermitianSymmetrizeX
  conv3  tests/hybridconv3.cc:57-64  f_a[i][j][k] = (i+ak+1) + i((a+1)j+3+k)
and the expected outputs restate the reference's direct sums:
  directconv::convolve / convolveC (tests/direct.h:17-39),
  directconvh::convolve (tests/direct.cc:5-14),
  directconv2::convolve (tests/direct.h:77-88),
  directconvh2::convolve (tests/direct.cc:16-50),
  directconv3::convolve (tests/direct.h), with HermitianSymmetrizeX as in
  convolve.h:1175-1197.
Complex numbers are (re, im) pairs of Python ints.

Run: python3 tests/golden/make_golden.py   (rewrites tests/golden/golden.json)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def mul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def add(a, b):
    return (a[0] + b[0], a[1] + b[1])


def conj(a):
    return (a[0], -a[1])


Z = (0, 0)


def ramp1(L, a):
    return [(j + a + 1, (a + 1) * j + 3) for j in range(L)]


def conv1(f, g):
    m = len(f)
    h = []
    for i in range(m):
        s = Z
        for j in range(i + 1):
            s = add(s, mul(f[j], g[i - j]))
        h.append(s)
    return h


def convC1(f, g):
    m = len(f)
    h = [Z] * m
    for i in range(m // 2):
        s = Z
        for j in range(i + 1, m):
            s = add(s, mul(f[j], g[m + i - j]))
        h[i + (m + 1) // 2] = s
    for i in range(m // 2, m):
        s = Z
        for j in range(i + 1):
            s = add(s, mul(f[j], g[i - j]))
        h[i - m // 2] = s
    return h


def convh1(f, g):
    m = len(f)
    h = []
    for i in range(m):
        s = Z
        for j in range(i + 1):
            s = add(s, mul(f[j], g[i - j]))
        for j in range(i + 1, m):
            s = add(s, mul(f[j], conj(g[j - i])))
        for j in range(1, m - i):
            s = add(s, mul(conj(f[j]), g[i + j]))
        h.append(s)
    return h


def ramp2(Lx, Ly, a):
    return [[(i + a + 1, (a + 1) * j + 3) for j in range(Ly)] for i in range(Lx)]


def conv2(f, g):
    mx, my = len(f), len(f[0])
    h = [[Z] * my for _ in range(mx)]
    for i in range(mx):
        for j in range(my):
            s = Z
            for k in range(i + 1):
                for p in range(j + 1):
                    s = add(s, mul(f[k][p], g[i - k][j - p]))
            h[i][j] = s
    return h


def rampH2(Lx, Hy, a):
    f = []
    for i in range(Lx):
        I = i if Lx % 2 else i - 1
        f.append([(I + a + 1, (a + 1) * j + 3) for j in range(Hy)])
    return f


def symmetrize_x(f, Hx, x0):
    # HermitianSymmetrizeX(Hx, Hy, x0, f): f[x0 - i][0] = conj f[x0 + i][0]
    # for 0 < i < Hx, Im f[x0][0] = 0, and with x0 == Hx the first row zeroed
    for i in range(1, Hx):
        f[x0 - i][0] = conj(f[x0 + i][0])
    f[x0][0] = (f[x0][0][0], 0)
    if x0 == Hx:
        f[0] = [Z] * len(f[0])


def convh2(f, g, Lx):
    Hy = len(f[0])
    mx = (Lx + 1) // 2
    xcompact = Lx % 2
    xorigin = mx - xcompact
    xstart, ystart, xstop, ystop = -xorigin, 1 - Hy, mx, Hy
    lo = xstart + (0 if xcompact else 1)

    def at(a, px, py):
        return a[xorigin + px][py] if py >= 0 else conj(a[xorigin - px][-py])
    h = [[Z] * Hy for _ in range(Lx)]
    for kx in range(xstart, xstop):
        for ky in range(ystop):
            s = Z
            for px in range(lo, xstop):
                qx = kx - px
                if not (lo <= qx < xstop):
                    continue
                for py in range(ystart, ystop):
                    qy = ky - py
                    if ystart <= qy < ystop:
                        s = add(s, mul(at(f, px, py), at(g, qx, qy)))
            h[xorigin + kx][ky] = s
    return h


def ramp3(Lx, Ly, Lz, a):
    return [[[(i + a * k + 1, (a + 1) * j + 3 + k) for k in range(Lz)] for j in range(Ly)]
            for i in range(Lx)]


def conv3(f, g):
    mx, my, mz = len(f), len(f[0]), len(f[0][0])
    h = [[[Z] * mz for _ in range(my)] for _ in range(mx)]
    for i in range(mx):
        for j in range(my):
            for k in range(mz):
                s = Z
                for a in range(i + 1):
                    for b in range(j + 1):
                        for c in range(k + 1):
                            s = add(s, mul(f[a][b][c], g[i - a][j - b][k - c]))
                h[i][j][k] = s
    return h


def flat(x):
    if isinstance(x, tuple):
        return [list(x)]
    out = []
    for v in x:
        out += flat(v)
    return out


def main():
    cases = []
    for L in [1, 2, 7, 8, 20, 64]:
        f, g = ramp1(L, 0), ramp1(L, 1)
        cases.append({"kind": "conv1", "shape": [L], "f": flat(f), "g": flat(g), "h": flat(conv1(f, g))})
    for L in [1, 7, 8, 20]:
        f, g = ramp1(L, 0), ramp1(L, 1)
        cases.append({"kind": "convC1", "shape": [L], "f": flat(f), "g": flat(g), "h": flat(convC1(f, g))})
    for H in [1, 2, 7, 8, 20]:
        f, g = ramp1(H, 0), ramp1(H, 1)
        f[0] = (f[0][0], 0)
        g[0] = (g[0][0], 0)
        cases.append({"kind": "convh1", "shape": [H], "L": 2 * H - 1, "f": flat(f), "g": flat(g),
                      "h": flat(convh1(f, g))})
    for Lx, Ly in [(7, 8), (8, 7), (1, 5), (5, 1)]:
        f, g = ramp2(Lx, Ly, 0), ramp2(Lx, Ly, 1)
        cases.append({"kind": "conv2", "shape": [Lx, Ly], "f": flat(f), "g": flat(g), "h": flat(conv2(f, g))})
    for Lx, Hy in [(7, 4), (9, 5), (8, 4), (3, 1)]:
        f, g = rampH2(Lx, Hy, 0), rampH2(Lx, Hy, 1)
        Hx = (Lx + 1) // 2
        for a in (f, g):
            symmetrize_x(a, Hx, Lx // 2)
        cases.append({"kind": "convh2", "shape": [Lx, Hy], "L": [Lx, 2 * Hy - 1], "f": flat(f),
                      "g": flat(g), "h": flat(convh2(f, g, Lx))})
    for Lx, Ly, Lz in [(4, 5, 3), (3, 2, 6)]:
        f, g = ramp3(Lx, Ly, Lz, 0), ramp3(Lx, Ly, Lz, 1)
        cases.append({"kind": "conv3", "shape": [Lx, Ly, Lz], "f": flat(f), "g": flat(g),
                      "h": flat(conv3(f, g))})
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump({"source": "tests/golden/make_golden.py (exact integer arithmetic)",
                   "cases": cases}, fh, separators=(",", ":"))
        fh.write("\n")


if __name__ == "__main__":
    main()

"""Writes tk_galerkin_pattern.hpp: the non-zero pattern of the Galerkin coarse
element operator in the Walsh-Hadamard sector basis, and the straight-line
device code that applies it.

    python pytopo3d_amd/csrc/gen_galerkin.py

A coarse element of the Galerkin level is the union of 8 child elements of the
level above, each with its own scale s_c.  Its operator is
K_E = sum_c s_c P_c^T KE P_c (P_c: trilinear interpolation from the coarse
corners to child c's corners).  In the kernel's Walsh-Hadamard basis (corner
modes m with bit 0 = x, bit 1 = y, bit 2 = z; entries 3 m + component, scaled
by 1/64 as the sector entries) this is sum_w shat_w G_w with shat the Walsh
transform of the 8 child scales over the child positions, and G_w couples the
parity sector chi only with chi XOR w.  The pattern is the same for every
axis-aligned orthotropic element (checked here on three KEs); the plan checks
its own KE against it at creation.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pytopo3d_amd.problem import lk_H8  # noqa: E402


def corner(a):
    return ((0xCC >> a) & 1, (0x66 >> a) & 1, (0xF0 >> a) & 1)   # (x, y, z)


def walsh_matrix():
    W = np.zeros((24, 24))
    for a in range(8):
        o = corner(a)
        for m in range(8):
            v = 1.0
            for d in range(3):
                if (m >> d) & 1:
                    v *= 1.0 if o[d] else -1.0
            for k in range(3):
                W[3 * a + k, 3 * m + k] = v
    return W


def child_interp(c):
    """24x24: child c's corner DOFs from the coarse element's corner DOFs."""
    cx, cy, cz = c & 1, (c >> 1) & 1, c >> 2

    def w1(f, b):
        return 1.0 if f == 2 * b else (0.5 if f == 1 else 0.0)

    P = np.zeros((24, 24))
    for a in range(8):
        ax, ay, az = corner(a)
        for b in range(8):
            bx, by, bz = corner(b)
            w = w1(cx + ax, bx) * w1(cy + ay, by) * w1(cz + az, bz)
            for k in range(3):
                P[3 * a + k, 3 * b + k] = w
    return P


def galerkin_blocks(KE):
    W = walsh_matrix()
    B = [W.T @ (child_interp(c).T @ KE @ child_interp(c)) @ W / 64.0 for c in range(8)]
    G = np.zeros((8, 24, 24))
    for w in range(8):
        for c in range(8):
            G[w] += (-1) ** bin(w & c).count("1") * B[c] / 8.0
    return G


def pattern():
    kes = [lk_H8(), lk_H8(10.0, 20.0, 30.0, 4.0, 5.0, 6.0, 0.1, 0.12, 0.08),
           lk_H8(1.0, 1.0, 1.0, 0.4, 0.4, 0.4, 0.3, 0.3, 0.3, elem_size=0.5)]
    pats = []
    for KE in kes:
        G = galerkin_blocks(KE)
        tol = 1e-12 * np.abs(G).max()
        pats.append({(w, i, j) for w in range(8) for i in range(24) for j in range(24)
                     if abs(G[w, i, j]) > tol})
    assert pats[0] == pats[1] == pats[2], "pattern differs between orthotropic KEs"
    pat = sorted(pats[0])
    # translations (mode 0 of each component) are annihilated
    assert all(i >= 3 and j >= 3 for _, i, j in pat)
    return pat


def main():
    pat = pattern()
    groups = {}
    for w, i, j in pat:
        groups.setdefault((w, i), []).append(j)
    keys = sorted(groups)
    ws, rows, cols, start = [], [], [], [0]
    for key in keys:
        ws.append(key[0])
        rows.append(key[1])
        cols.extend(groups[key])
        start.append(len(cols))
    lines = [
        "// Generated by gen_galerkin.py: the Galerkin coarse element operator in the",
        "// Walsh-Hadamard sector basis (see that script).  Do not edit.",
        "#pragma once",
        "",
        '#include "tk_common.hpp"',
        "",
        "namespace tk {",
        "",
        f"constexpr int kGalN = {len(cols)};      // non-zero coupling entries",
        f"constexpr int kGalGroups = {len(keys)};  // (Walsh mode, output row) groups",
        "// group g: Walsh mode kGalW[g] of the child scales, output entry kGalRow[g]",
        "// (= 3 corner mode + component), table entries kGalStart[g] .. kGalStart[g + 1] - 1",
        "// with input entries kGalCol[k]",
        "constexpr unsigned char kGalW[kGalGroups] = {" + ", ".join(map(str, ws)) + "};",
        "constexpr unsigned char kGalRow[kGalGroups] = {" + ", ".join(map(str, rows)) + "};",
        "constexpr short kGalStart[kGalGroups + 1] = {" + ", ".join(map(str, start)) + "};",
        "constexpr unsigned char kGalCol[kGalN] = {" + ", ".join(map(str, cols)) + "};",
        "",
        "// o = sum_w sh[w] G_w a (spectra a[c][m], o[c][m], mode 0 unused and o[c][0] = 0);",
        "// gk: the plan's table in the order above.",
        "template <typename T>",
        "__device__ __forceinline__ void galerkin_apply(const T* gk, const T sh[8], const T a[3][8],",
        "                                               T o[3][8]) {",
        "#pragma unroll",
        "  for (int c = 0; c < 3; ++c)",
        "#pragma unroll",
        "    for (int m = 0; m < 8; ++m) o[c][m] = T(0);",
    ]
    for g, key in enumerate(keys):
        w, i = key
        terms = " + ".join(f"gk[{k}] * a[{cols[k] % 3}][{cols[k] // 3}]"
                           for k in range(start[g], start[g + 1]))
        lines.append(f"  o[{i % 3}][{i // 3}] += sh[{w}] * ({terms});")
    lines += ["}", "", "}  // namespace tk", ""]
    out = os.path.join(HERE, "tk_galerkin_pattern.hpp")
    with open(out, "w") as fh:
        fh.write("\n".join(lines))
    print(f"wrote {out}: {len(cols)} entries in {len(keys)} groups")


if __name__ == "__main__":
    main()

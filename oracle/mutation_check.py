"""Mutation check of the oracle's pinning tests (test infrastructure).

Plants one plausible mistake at a time in a copy of rgl_oracle.cpp, builds
the copy into a temporary directory and runs the CPU pinning tests against
it (RGL_ORACLE_SO); every mistake must make a test fail.  The oracle's own
source is never modified.

  python oracle/mutation_check.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (what the mistake is, the line as written, the line with the mistake)
MUTANTS = [
    ("vector top-left bias sign", "return (dy > 0 || (dy == 0 && dx > 0)) ? 1 : 0; };",
     "return (dy > 0 || (dy == 0 && dx < 0)) ? 1 : 0; };"),
    ("vector quad lane order", "int32_t e1[4] = { q1, wadd(q1, dy12), wadd(q1, dx12), wadd(wadd(q1, dx12), dy12) };",
     "int32_t e1[4] = { q1, wadd(q1, dx12), wadd(q1, dy12), wadd(wadd(q1, dx12), dy12) };"),
    ("perspective-correct w index", "bp[0][l] = (tc.w[0] * bs[0][l]) * fragW;", "bp[0][l] = (tc.w[2] * bs[0][l]) * fragW;"),
    ("no backface reorder", "uint16_t n0 = front ? ii[0] : ii[2];", "uint16_t n0 = ii[0];"),
    ("LEQUAL tested as LESS", "case GL_LEQUAL: return fragDepth[l] <= dst;", "case GL_LEQUAL: return fragDepth[l] < dst;"),
    ("bbox max without the +1", "const int32_t vmaxx = std::min(wadd(std::max(ix0, std::max(ix1, ix2)), 1), brx);",
     "const int32_t vmaxx = std::min(std::max(ix0, std::max(ix1, ix2)), brx);"),
    ("primitive ids triangle-major", "const uint32_t prim = primBase + static_cast<uint32_t>(iid) * numPrims + pi;",
     "const uint32_t prim = primBase + static_cast<uint32_t>(pi) * instanceCnt + iid;"),
    ("clip interpolation not inside->outside", "const ClippedVertex& from = hereIsInside ? A[h] : A[n];",
     "const ClippedVertex& from = A[h];"),
    ("clip plane sign", "case 1: return p[3] + p[1];", "case 1: return p[3] - p[1];"),
    ("nearest sampler v not flipped", "const int32_t ly = cvtt((almostone - frv) * levelDim);",
     "const int32_t ly = cvtt(frv * levelDim);"),
    ("npot sampler row stride", "const int32_t ofs = shl(wadd(wmul(pv, stride), pu), 2);",
     "const int32_t ofs = shl(wadd(wmul(pv, width), pu), 2);"),
    ("bilinear half-texel offset", "const float fx = (lx - itof(tx0)) - 0.5F;", "const float fx = (lx - itof(tx0));"),
    ("GGX D without pi", "const float D = a2 / ((kPi * t) * t);", "const float D = a2 / (t * t);"),
    ("GGX G(n.v) dropped", "const float spec = (((D * gl) * gv) * F) / ((4.0F * ndl) * ndv);",
     "const float spec = (((D * gl)) * F) / ((4.0F * ndl) * ndv);"),
    ("environment face order", "face = dx > 0.0F ? 0 : 1;", "face = dx > 0.0F ? 1 : 0;"),
    ("Blinn-Phong exponent 16", "float s = ndh * ndh; s = s * s; s = s * s; s = s * s; s = s * s;",
     "float s = ndh * ndh; s = s * s; s = s * s; s = s * s;"),
    ("eye light difference with w from p.w", "lightPos[2] - p[2], lightPos[3] - p[2] };",
     "lightPos[2] - p[2], lightPos[3] - p[3] };"),
    ("clipping rasteriser gl_FragCoord row", "fc[1][l] = fy + static_cast<float>(l < 2 ? 1 : 0);",
     "fc[1][l] = fy - static_cast<float>(l >> 1);"),
    ("alpha discard keeps alpha 0", "for (int l = 0; l < 4; ++l) mask[l] &= (color[3][l] > 0.0F) ? -1 : 0; }};",
     "for (int l = 0; l < 4; ++l) mask[l] &= (color[3][l] >= 0.0F) ? -1 : 0; }};"),
    ("depth unit v not flipped", "const int32_t py = cvtt((1.0F - y[l]) * dimf);", "const int32_t py = cvtt(y[l] * dimf);"),
    ("normal matrix not transposed", "mat4_transpose(inv.ff, env.nm.ff);", "env.nm = inv;"),
    ("spot cone test inverted", "if (ang < u[22]) att[l] = 0.111F;", "if (ang > u[22]) att[l] = 0.111F;"),
    ("text colour slots swapped", "o.f[0] = b[6][idx]; o.f[1] = b[7][idx]; o.f[2] = b[8][idx];",
     "o.f[0] = b[8][idx]; o.f[1] = b[7][idx]; o.f[2] = b[6][idx];"),
    ("pattern divided by dim.x", "uu[l] = fc[0][l] / u[5]; vv[l] = fc[1][l] / u[5];",
     "uu[l] = fc[0][l] / u[4]; vv[l] = fc[1][l] / u[4];"),
    ("envmap reflection sign", "for (int k = 0; k < 3; ++k) r[k] = ev[k] - twod * n[k];",
     "for (int k = 0; k < 3; ++k) r[k] = ev[k] + twod * n[k];"),
    ("exp2 integer part truncated", "const int32_t ipart = cvtn(x - 0.5F);", "const int32_t ipart = cvtt(x - 0.5F);"),
    ("IQ gamma of red and blue swapped", "c[0] = sse_pow(c[0], 0.45F); c[1] = sse_pow(c[1], 0.5F); c[2] = sse_pow(c[2], 0.55F);",
     "c[0] = sse_pow(c[0], 0.55F); c[1] = sse_pow(c[1], 0.5F); c[2] = sse_pow(c[2], 0.45F);"),
    ("IQ vignette exponent", "const float v = 0.2F + 0.8F * sse_pow(t, 0.2F);", "const float v = 0.2F + 0.8F * sse_pow(t, 0.25F);"),
]


def main() -> int:
    src = open(os.path.join(HERE, "rgl_oracle.cpp")).read()
    missed = 0
    with tempfile.TemporaryDirectory() as tmp:
        for inc in ("rcpps_table.inc", "rsqrtps_table.inc", "srgb8_table.inc"):
            os.symlink(os.path.join(HERE, inc), os.path.join(tmp, inc))
        for name, good, bad in MUTANTS:
            assert src.count(good) == 1, name
            cpp = os.path.join(tmp, "rgl_oracle.cpp")
            so = os.path.join(tmp, f"m{MUTANTS.index((name, good, bad))}.so")
            open(cpp, "w").write(src.replace(good, bad))
            subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-shared", "-o", so, cpp,
                            "-lpthread"], check=True)
            env = dict(os.environ, RGL_ORACLE_SO=so)
            r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "tests/test_oracle_pins.py",
                                "tests/test_program_pins.py", "tests/test_oracle_cpu.py", "-k", "not host_hardware and not reference_header and not user_programs"],
                               cwd=ROOT, env=env, capture_output=True, text=True)
            caught = r.returncode != 0
            failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
            print(f"{'caught' if caught else 'MISSED'}: {name}" + (f"  ({failed[0][7:80]})" if failed else ""), flush=True)
            missed += 0 if caught else 1
    print(f"{len(MUTANTS) - missed} of {len(MUTANTS)} planted mistakes caught")
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main())

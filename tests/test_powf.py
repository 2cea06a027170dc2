"""The Poisson term's pow shortcut (gpuvoxels_amd/csrc/vx3_powf.h) against the value the
reference computes, pow of two floats evaluated in double and rounded once
(Voxelyze3/src/old/VX_Voxel.cpp:299-333): whenever the shortcut answers it must give
the same float, bit for bit.  The header is host-compilable; this builds it with g++."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include "vx3_powf.h"
int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<float> wide(-0.3f, 0.3f), nus(0.0f, 0.499f);
    long answered = 0, mismatches = 0, declined = 0;
    for (long i = 0; i < 4000000; i++) {
        float add = wide(g);
        if (i % 3 == 0) add *= 1e-3f;                 // the common case: small strains
        if (i % 97 == 0) add = ldexpf(1.0f, -(int)(i % 40));
        const float nu = (i % 4) ? 0.35f : nus(g);
        const float base = 1.0f + add;
        const float ref = (float)pow((double)base, (double)(-nu));
        float f;
        if (vx3::powNearOne(base, -nu, f)) {
            answered++;
            if (memcmp(&f, &ref, 4) != 0) mismatches++;
        } else declined++;
    }
    // the boundaries: exactly 1, out of range, NaN
    float f;
    const bool one = vx3::powNearOne(1.0f, -0.35f, f) && f == 1.0f;
    const bool far = !vx3::powNearOne(1.5f, -0.35f, f) && !vx3::powNearOne(0.5f, -0.35f, f);
    const bool nan = !vx3::powNearOne(NAN, -0.35f, f);
    printf("%ld %ld %ld %d %d %d\n", answered, mismatches, declined, (int)one, (int)far, (int)nan);
    return 0;
}
"""


def test_pow_shortcut_equals_rounded_double_pow(tmp_path):
    src = tmp_path / "powf.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "powf"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "gpuvoxels_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True, timeout=120).stdout.split()
    answered, mismatches, declined, one, far, nan = map(int, out)
    assert mismatches == 0
    assert answered > 0.8 * (answered + declined)      # declines only outside |base-1| <= 1/4 or near ties
    assert one == 1 and far == 1 and nan == 1

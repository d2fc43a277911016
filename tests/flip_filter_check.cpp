// Checks flip_decide (pamo_amd/csrc/flip_filter.h) against the reference's
// normalise-then-dot flip test (simp_cuda/src/cusimp_free.cu:520-526) on
// random, near-perpendicular, wide-exponent and special inputs.  Built with
// g++ -ffp-contract=off by tests/test_flip_filter.py; prints
//   <case> total decided mismatches
// per case, and exits non-zero on any mismatch.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "flip_filter.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}
static float unif(float lo, float hi) { return lo + (hi - lo) * (float)((next_u64() >> 40) * (1.0 / 16777216.0)); }
static float pow2f(int e) { return ldexpf(1.0f, e); }

// the reference's expression, FP32, in its order
static int flip_ref(float ax, float ay, float az, float bx, float by, float bz) {
  const float L = sqrtf(bx * bx + by * by + bz * bz);
  const float qx = bx / L, qy = by / L, qz = bz / L;
  return ax * qx + ay * qy + az * qz < 0.0f;
}

struct Tally { long long total = 0, decided = 0, bad = 0; };

static void check(Tally& t, float ax, float ay, float az, float bx, float by, float bz) {
  const int d = pamo::flipf::flip_decide(ax, ay, az, bx, by, bz);
  t.total++;
  if (d < 0) return;
  t.decided++;
  if (d != flip_ref(ax, ay, az, bx, by, bz)) {
    if (t.bad < 5)
      fprintf(stderr, "mismatch a=(%a %a %a) b=(%a %a %a) filter=%d ref=%d\n", ax, ay, az, bx, by, bz, d,
              flip_ref(ax, ay, az, bx, by, bz));
    t.bad++;
  }
}

// a unit normal as the reference forms one: c / |c| in FP32
static void unit(float& x, float& y, float& z) {
  float cx, cy, cz, n;
  do {
    cx = unif(-1, 1); cy = unif(-1, 1); cz = unif(-1, 1);
    n = sqrtf(cx * cx + cy * cy + cz * cz);
  } while (!(n > 0.05f));
  x = cx / n; y = cy / n; z = cz / n;
}

int main(int argc, char** argv) {
  long long n = argc > 1 ? atoll(argv[1]) : 2000000;
  int fails = 0;
  auto report = [&](const char* name, const Tally& t) {
    printf("%s %lld %lld %lld\n", name, t.total, t.decided, t.bad);
    if (t.bad) fails++;
  };
  {  // random directions, magnitudes 2^-30 .. 2^30
    Tally t;
    for (long long i = 0; i < n; i++) {
      float ax, ay, az, bx, by, bz;
      unit(ax, ay, az);
      unit(bx, by, bz);
      const float s = pow2f((int)(next_u64() % 61) - 30) * unif(1, 2);
      check(t, ax, ay, az, bx * s, by * s, bz * s);
    }
    report("random", t);
  }
  {  // b nearly perpendicular to a: b = w - (w.a) a + eps a, eps = 2^-k
    Tally t;
    for (long long i = 0; i < n; i++) {
      float ax, ay, az, wx, wy, wz;
      unit(ax, ay, az);
      unit(wx, wy, wz);
      const float wa = wx * ax + wy * ay + wz * az;
      const float eps = pow2f(-(int)(next_u64() % 40)) * unif(-1, 1);
      const float s = pow2f((int)(next_u64() % 61) - 30);
      const float bx = (wx - wa * ax + eps * ax) * s, by = (wy - wa * ay + eps * ay) * s,
                  bz = (wz - wa * az + eps * az) * s;
      check(t, ax, ay, az, bx, by, bz);
    }
    report("near_perpendicular", t);
  }
  {  // raw components with random exponents and signs, |a_i| up to 4
    Tally t;
    for (long long i = 0; i < n; i++) {
      float v[6];
      for (int j = 0; j < 6; j++) {
        const int e = j < 3 ? -(int)(next_u64() % 30) + 2 : (int)(next_u64() % 200) - 150;
        v[j] = ((next_u64() & 1) ? -1.0f : 1.0f) * pow2f(e) * unif(1, 2);
        if (next_u64() % 8 == 0) v[j] = 0.0f;
      }
      check(t, v[0], v[1], v[2], v[3], v[4], v[5]);
    }
    report("wide_exponent", t);
  }
  {  // specials: zeros, signed zeros, denormals, infinities, NaN
    Tally t;
    const float sp[] = {0.0f, -0.0f, 1e-45f, -1e-45f, 1e-39f, INFINITY, -INFINITY, NAN, 1.0f, -1.0f, 0x1p-40f,
                        0x1p40f, 0x1p-64f, 3.4e38f};
    const int ns = sizeof(sp) / sizeof(sp[0]);
    for (long long i = 0; i < n / 4; i++) {
      float v[6];
      for (int j = 0; j < 6; j++) v[j] = (next_u64() % 2) ? sp[next_u64() % ns] : unif(-1, 1);
      check(t, v[0], v[1], v[2], v[3], v[4], v[5]);
    }
    report("specials", t);
  }
  return fails ? 1 : 0;
}

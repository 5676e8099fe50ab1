/*
 * Float64 C restatement of the oracle step (TEST INFRASTRUCTURE ONLY).
 *
 * Same algorithm as oracle/lbm_ref.py (pylbm/core.py:88-247): pull streaming where a link
 * that crosses any wall bounces back the node's own opposite population (plus Ladd's term
 * 6 w_i c_ix U when the crossing is through the top wall of the last axis), then moments
 * with the Guo half-force, equilibrium, and BGK / TRT / MRT collision with Guo forcing.
 * Written per node so it runs the 64^3 x 500-step checks in seconds; pinned against the
 * NumPy oracle in tests/test_oracle_c.py.  Only tests and the CPU-baseline leg of bench.py
 * load it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int C2[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0},
                             {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
static const int OPP2[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const int C3[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                              {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
                              {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                              {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
static const int MM[9][9] = {{1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2},
                             {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
                             {0, -2, 0, 2, 0, 1, -1, -1, 1},  {0, 0, 1, 0, -1, 1, 1, -1, -1},
                             {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},
                             {0, 0, 0, 0, 0, 1, -1, 1, -1}};

typedef struct {
  int lattice;        /* 0 D2Q9, 1 D3Q19 */
  int collision;      /* 0 BGK, 1 TRT, 2 MRT */
  int periodic[3];
  int nx, ny, nz;
  double ulid;
  double force[3];
  const double* force_field; /* (d, N) or NULL */
  double rates[9];    /* BGK: omega; TRT: s+, s-; MRT: the 9 diagonal rates */
} oracle_case;

static void collide(const oracle_case* cs, int q, int d, const int (*c)[3], const int* opp,
                    const double* w, const double* f, const double F[3], int has_force,
                    double* p) {
  double rho = f[0];
  for (int i = 1; i < q; ++i) rho += f[i];
  double u[3] = {0, 0, 0};
  for (int a = 0; a < d; ++a) {
    double j = 0;
    for (int i = 0; i < q; ++i) j += c[i][a] * f[i];
    u[a] = (j + 0.5 * F[a]) / rho;
  }
  double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  double uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
  double feq[19], S[19];
  for (int i = 0; i < q; ++i) {
    double cu = c[i][0] * u[0] + c[i][1] * u[1] + c[i][2] * u[2];
    feq[i] = w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    double cf = c[i][0] * F[0] + c[i][1] * F[1] + c[i][2] * F[2];
    S[i] = has_force ? w[i] * (3.0 * (cf - uf) + 9.0 * cu * cf) : 0.0;
  }
  const double* r = cs->rates;
  if (cs->collision == 0) {
    for (int i = 0; i < q; ++i)
      p[i] = f[i] - r[0] * (f[i] - feq[i]) + (1.0 - 0.5 * r[0]) * S[i];
  } else if (cs->collision == 1) {
    for (int i = 0; i < q; ++i) {
      int o = opp[i];
      p[i] = f[i] - r[0] * (0.5 * (f[i] + f[o]) - 0.5 * (feq[i] + feq[o])) -
             r[1] * (0.5 * (f[i] - f[o]) - 0.5 * (feq[i] - feq[o])) +
             (1.0 - 0.5 * r[0]) * 0.5 * (S[i] + S[o]) + (1.0 - 0.5 * r[1]) * 0.5 * (S[i] - S[o]);
    }
  } else {
    double dm[9];
    for (int k = 0; k < 9; ++k) {
      double a = 0, b = 0, nrm = 0;
      for (int i = 0; i < 9; ++i) {
        a += MM[k][i] * (f[i] - feq[i]);
        b += MM[k][i] * S[i];
        nrm += MM[k][i] * MM[k][i];
      }
      dm[k] = (-r[k] * a + (1.0 - 0.5 * r[k]) * b) / nrm;
    }
    for (int i = 0; i < 9; ++i) {
      double v = f[i];
      for (int k = 0; k < 9; ++k) v += MM[k][i] * dm[k];
      p[i] = v;
    }
  }
}

/* Advance `steps` steps of the canonical float64 state f (q * N values, in place). */
int oracle_run(const oracle_case* cs, int64_t steps, double* f) {
  const int d = cs->lattice == 0 ? 2 : 3, q = cs->lattice == 0 ? 9 : 19;
  const int(*c)[3] = cs->lattice == 0 ? C2 : C3;
  int opp[19];
  double w[19];
  for (int i = 0; i < q; ++i) {
    opp[i] = cs->lattice == 0 ? OPP2[i] : (i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1));
    int s = abs(c[i][0]) + abs(c[i][1]) + abs(c[i][2]);
    w[i] = cs->lattice == 0 ? (s == 0 ? 4.0 / 9 : (s == 1 ? 1.0 / 9 : 1.0 / 36))
                            : (s == 0 ? 1.0 / 3 : (s == 1 ? 1.0 / 18 : 1.0 / 36));
  }
  const int nx = cs->nx, ny = cs->ny, nz = d == 3 ? cs->nz : 1;
  const int64_t N = (int64_t)nx * ny * nz;
  const int dims[3] = {nx, ny, nz};
  double* g = (double*)malloc(sizeof(double) * q * N);
  if (!g) return 1;
  const int has_force = cs->force_field || cs->force[0] != 0 || cs->force[1] != 0 ||
                        cs->force[2] != 0;
  for (int64_t t = 0; t < steps; ++t) {
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
      const int x[3] = {(int)(n % nx), (int)((n / nx) % ny), (int)(n / ((int64_t)nx * ny))};
      double fl[19], p[19];
      for (int i = 0; i < q; ++i) {
        int s[3], wall = 0, top = 0;
        for (int a = 0; a < 3; ++a) {
          s[a] = x[a] - c[i][a];
          if (a < d && (s[a] < 0 || s[a] >= dims[a])) {
            if (cs->periodic[a]) s[a] = (s[a] + dims[a]) % dims[a];
            else {
              wall = 1;
              if (a == d - 1 && s[a] >= dims[a]) top = 1;
            }
          }
        }
        if (wall) {
          fl[i] = f[(int64_t)opp[i] * N + n];
          if (top && cs->ulid != 0.0) fl[i] += 6.0 * w[i] * c[i][0] * cs->ulid;
        } else {
          fl[i] = f[(int64_t)i * N + s[0] + (int64_t)nx * (s[1] + (int64_t)ny * s[2])];
        }
      }
      double F[3] = {cs->force[0], cs->force[1], d == 3 ? cs->force[2] : 0.0};
      if (cs->force_field)
        for (int a = 0; a < d; ++a) F[a] = cs->force_field[(int64_t)a * N + n];
      collide(cs, q, d, c, opp, w, fl, F, has_force, p);
      for (int i = 0; i < q; ++i) g[(int64_t)i * N + n] = p[i];
    }
    memcpy(f, g, sizeof(double) * q * N);
  }
  free(g);
  return 0;
}

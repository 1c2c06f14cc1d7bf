// Shape functions, Gauss rules and the hyperelastic constitutive update, shared
// by the element, mass and adjoint kernels.
#pragma once

#include "common.h"

namespace fea {

// ------------------------------------------------------------------ elements
// Reference shape functions / Gauss rules:
//   C3D4  tetrahedral.py:30-40  (N = 1-g-h-r, g, h, r; 1 GP (1/4,1/4,1/4), w 1/6)
//   C3D8  brick.py:52-160        (trilinear, 2x2x2 GP +-1/sqrt3 in brick order, w 1)
//   C3D10 tetrahedral.py:94-116  (quadratic, 4 GP with the truncated constants, w 1/24)
template <int ET>
struct Elem;

template <>
struct Elem<4> {
  static constexpr int nen = 4, ngp = 1;
  __device__ static constexpr double weight(int) { return 1.0 / 6.0; }
  __device__ static void shape(int, double N[4]) {
    N[0] = 0.25; N[1] = 0.25; N[2] = 0.25; N[3] = 0.25;
  }
  // structural zeros of dN[m][a] (skipped in the sums below: x * 0 is not folded away)
  __device__ static constexpr bool dzero(int m, int a) { return a != 0 && a != m + 1; }
  // dN[m][a] = dN_a / dxi_m
  __device__ static void dshape(int, double dN[3][4]) {
    dN[0][0] = -1.0; dN[0][1] = 1.0; dN[0][2] = 0.0; dN[0][3] = 0.0;
    dN[1][0] = -1.0; dN[1][1] = 0.0; dN[1][2] = 1.0; dN[1][3] = 0.0;
    dN[2][0] = -1.0; dN[2][1] = 0.0; dN[2][2] = 0.0; dN[2][3] = 1.0;
  }
};

__device__ __forceinline__ double hex_sign(int a, int axis) {
  // node / Gauss-point sign pattern of brick.py:13-22 and :149-158
  const int gs = ((a == 1) || (a == 2) || (a == 5) || (a == 6)) ? 1 : -1;
  const int hs = ((a == 2) || (a == 3) || (a == 6) || (a == 7)) ? 1 : -1;
  const int rs = (a >= 4) ? 1 : -1;
  return (double)(axis == 0 ? gs : (axis == 1 ? hs : rs));
}

template <>
struct Elem<8> {
  static constexpr int nen = 8, ngp = 8;
  static constexpr double q = 0.57735026918962576451;  // 1/sqrt(3)
  __device__ static constexpr double weight(int) { return 1.0; }
  __device__ static constexpr bool dzero(int, int) { return false; }
  __device__ static void shape(int g, double N[8]) {
    const double x = q * hex_sign(g, 0), y = q * hex_sign(g, 1), z = q * hex_sign(g, 2);
#pragma unroll
    for (int a = 0; a < 8; ++a)
      N[a] = 0.125 * (1.0 + x * hex_sign(a, 0)) * (1.0 + y * hex_sign(a, 1)) * (1.0 + z * hex_sign(a, 2));
  }
  __device__ static void dshape(int g, double dN[3][8]) {
    const double x = q * hex_sign(g, 0), y = q * hex_sign(g, 1), z = q * hex_sign(g, 2);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double sx = hex_sign(a, 0), sy = hex_sign(a, 1), sz = hex_sign(a, 2);
      dN[0][a] = 0.125 * sx * (1.0 + y * sy) * (1.0 + z * sz);
      dN[1][a] = 0.125 * sy * (1.0 + x * sx) * (1.0 + z * sz);
      dN[2][a] = 0.125 * sz * (1.0 + x * sx) * (1.0 + y * sy);
    }
  }
};

template <>
struct Elem<10> {
  static constexpr int nen = 10, ngp = 4;
  static constexpr double alpha = 0.58541020, beta = 0.13819660;  // tetrahedral.py:112-113
  __device__ static constexpr double weight(int) { return 1.0 / 24.0; }
  __device__ static constexpr bool dzero(int m, int a) {
    return (a == 1 && m != 0) || (a == 2 && m != 1) || (a == 3 && m != 2) || (a == 5 && m == 2) ||
           (a == 8 && m == 1) || (a == 9 && m == 0);
  }
  __device__ static void point(int g, double &x, double &y, double &z) {
    x = (g == 1) ? alpha : beta;
    y = (g == 2) ? alpha : beta;
    z = (g == 3) ? alpha : beta;
  }
  __device__ static void shape(int g, double N[10]) {
    double x, y, z;
    point(g, x, y, z);
    const double L = 1.0 - x - y - z;
    N[0] = L * (2.0 * L - 1.0);
    N[1] = x * (2.0 * x - 1.0);
    N[2] = y * (2.0 * y - 1.0);
    N[3] = z * (2.0 * z - 1.0);
    N[4] = 4.0 * x * L;
    N[5] = 4.0 * x * y;
    N[6] = 4.0 * y * L;
    N[7] = 4.0 * z * L;
    N[8] = 4.0 * x * z;
    N[9] = 4.0 * y * z;
  }
  __device__ static void dshape(int g, double dN[3][10]) {
    double x, y, z;
    point(g, x, y, z);
    const double L = 1.0 - x - y - z;
    const double d0 = 1.0 - 4.0 * L;  // dN0/dxi_m for every m
    dN[0][0] = d0; dN[1][0] = d0; dN[2][0] = d0;
    dN[0][1] = 4.0 * x - 1.0; dN[1][1] = 0.0; dN[2][1] = 0.0;
    dN[0][2] = 0.0; dN[1][2] = 4.0 * y - 1.0; dN[2][2] = 0.0;
    dN[0][3] = 0.0; dN[1][3] = 0.0; dN[2][3] = 4.0 * z - 1.0;
    dN[0][4] = 4.0 * (L - x); dN[1][4] = -4.0 * x; dN[2][4] = -4.0 * x;
    dN[0][5] = 4.0 * y; dN[1][5] = 4.0 * x; dN[2][5] = 0.0;
    dN[0][6] = -4.0 * y; dN[1][6] = 4.0 * (L - y); dN[2][6] = -4.0 * y;
    dN[0][7] = -4.0 * z; dN[1][7] = -4.0 * z; dN[2][7] = 4.0 * (L - z);
    dN[0][8] = 4.0 * z; dN[1][8] = 0.0; dN[2][8] = 4.0 * x;
    dN[0][9] = 0.0; dN[1][9] = 4.0 * z; dN[2][9] = 4.0 * y;
  }
};

// ------------------------------------------------------------------ material
// W = w1 (I1bar-3) + w2 (I2bar-3) + kv/2 (J-1)^2, isochoric invariants of
// C = F^T F.  NeoHookean (hyperelastic.py:25-33): w1 = mu/2, w2 = 0, kv = kappa.
struct MatParams {
  double w1, w2, kv;
};

// Voigt order of symmetric tensors: 00, 11, 22, 01, 12, 02
__device__ __forceinline__ constexpr int vA(int I) { return I < 3 ? I : (I == 3 ? 0 : (I == 4 ? 1 : 0)); }
__device__ __forceinline__ constexpr int vB(int I) { return I < 3 ? I : (I == 3 ? 1 : 2); }
__device__ __forceinline__ constexpr int vidx(int A, int B) {
  return A == B ? A : ((A + B == 1) ? 3 : ((A + B == 3) ? 4 : 5));
}

// Kinematics at one Gauss point.
template <typename T>
struct GPState {
  T F[9];         // F[i*3+A] = dx_i/dX_A
  T C[6];         // F^T F, Voigt
  T Ci[6];        // C^-1, Voigt
  T I1, I2;
  double J, Jm1;  // J and J-1 (float64, cancellation-free)
};

template <typename T>
__device__ __forceinline__ void kinematics(const T H[9], GPState<T> &s) {
  #pragma clang fp contract(on)
#pragma unroll
  for (int i = 0; i < 9; ++i) s.F[i] = H[i];
  s.F[0] += T(1); s.F[4] += T(1); s.F[8] += T(1);
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int A = vA(I), B = vB(I);
    s.C[I] = s.F[0 + A] * s.F[0 + B] + s.F[3 + A] * s.F[3 + B] + s.F[6 + A] * s.F[6 + B];
  }
  s.I1 = s.C[0] + s.C[1] + s.C[2];
  const T trC2 = s.C[0] * s.C[0] + s.C[1] * s.C[1] + s.C[2] * s.C[2] +
                 T(2) * (s.C[3] * s.C[3] + s.C[4] * s.C[4] + s.C[5] * s.C[5]);
  s.I2 = T(0.5) * (s.I1 * s.I1 - trC2);
  double hd[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) hd[i] = (double)H[i];
  s.Jm1 = detm1_of_grad(hd);
  s.J = 1.0 + s.Jm1;
  // C^-1 = adj(C) / det C, det C = J^2
  const T idet = T(1.0 / (s.J * s.J));
  const T c00 = s.C[0], c11 = s.C[1], c22 = s.C[2], c01 = s.C[3], c12 = s.C[4], c02 = s.C[5];
  s.Ci[0] = (c11 * c22 - c12 * c12) * idet;
  s.Ci[1] = (c00 * c22 - c02 * c02) * idet;
  s.Ci[2] = (c00 * c11 - c01 * c01) * idet;
  s.Ci[3] = (c02 * c12 - c01 * c22) * idet;
  s.Ci[4] = (c01 * c02 - c00 * c12) * idet;
  s.Ci[5] = (c01 * c12 - c02 * c11) * idet;
}

// J^(-2/3) for the stress of element code in T.  In double: pow.  In float: J = m 8^k with
// m in [1/2, 4) (frexp, exact), a float estimate of y = m^(-1/3) from the hardware log2 / exp2,
// refined by two float64 Newton steps on y^-3 = m, y <- y (4 - m y^3) / 3 (division-free; the
// relative error goes ~1e-7 -> ~1e-13 -> round-off), and J^(-2/3) = y^2 4^-k: the volumetric
// terms keep float64 accuracy over the whole double range for a fraction of a float64 pow's
// work and registers.  J <= 0 gives NaN (pow: NaN, or inf at 0).
template <typename T>
__device__ __forceinline__ double jm23(double J) {
  #pragma clang fp contract(on)
  if constexpr (sizeof(T) < sizeof(double)) {
    int e;
    const double mant = frexp(J, &e);
    const int k = e >= 0 ? e / 3 : -((2 - e) / 3);  // floor(e / 3)
    const double m = ldexp(mant, e - 3 * k);
    double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)m) * (-1.0f / 3.0f));
#pragma unroll
    for (int i = 0; i < 2; ++i) y = y * (4.0 - m * (y * y * y)) * (1.0 / 3.0);
    return ldexp(y * y, -2 * k);
  } else {
    return pow(J, -2.0 / 3.0);
  }
}

// Second Piola-Kirchhoff stress S = 2 dW/dC = s1 1 + sc C + si C^-1.
template <typename T>
__device__ __forceinline__ void stress(const GPState<T> &s, const MatParams &m, T S[6]) {
  #pragma clang fp contract(on)
  const double Jm23 = jm23<T>(s.J);
  const double Jm43 = Jm23 * Jm23;
  const double I1 = (double)s.I1, I2 = (double)s.I2;
  const double up = m.kv * s.Jm1;  // U'(J)
  const T s1 = T(2.0 * m.w1 * Jm23 + 2.0 * m.w2 * Jm43 * I1);
  const T sc = T(-2.0 * m.w2 * Jm43);
  const T si = T(-(2.0 / 3.0) * m.w1 * Jm23 * I1 - (4.0 / 3.0) * m.w2 * Jm43 * I2 + s.J * up);
#pragma unroll
  for (int I = 0; I < 6; ++I) S[I] = I < 3 ? sc * s.C[I] + si * s.Ci[I] + s1 : sc * s.C[I] + si * s.Ci[I];
}

// Material tangent CC = 2 dS/dC of W(I1bar, I2bar, J):
//   CC = cii Ci(x)Ci + c1i (1(x)Ci + Ci(x)1) + cci (C(x)Ci + Ci(x)C) + c11 1(x)1
//        + cI  Isym + cIi I_{Ci},   (I_{Ci})_ABCD = (Ci_AC Ci_BD + Ci_AD Ci_BC)/2
struct TangentCoeffs {
  double cii, c1i, cci, c11, cI, cIi;
};
template <typename T>
__device__ __forceinline__ TangentCoeffs tangent_coeffs(const GPState<T> &s, const MatParams &m) {
  const double Jm23 = pow(s.J, -2.0 / 3.0);
  const double Jm43 = Jm23 * Jm23;
  const double I1 = (double)s.I1, I2 = (double)s.I2;
  const double g1 = 4.0 * m.w1 * Jm23, g2 = 4.0 * m.w2 * Jm43;
  const double up = m.kv * s.Jm1, upp = m.kv;
  TangentCoeffs c;
  c.cii = g1 * I1 / 9.0 + (4.0 / 9.0) * g2 * I2 + s.J * (up + s.J * upp);
  c.c1i = -g1 / 3.0 - (2.0 / 3.0) * g2 * I1;
  c.cci = (2.0 / 3.0) * g2;
  c.c11 = g2;
  c.cI = -g2;
  c.cIi = g1 * I1 / 3.0 + (2.0 / 3.0) * g2 * I2 - 2.0 * s.J * up;
  return c;
}
// CC in Voigt form (engineering shear strain), full 6x6: D[I][J] = CC_{vA(I) vB(I) vA(J) vB(J)}.
template <typename T>
__device__ __forceinline__ void tangent(const GPState<T> &s, const MatParams &m, T D[6][6]) {
  const TangentCoeffs tc = tangent_coeffs(s, m);
  const T cii = T(tc.cii), c1i = T(tc.c1i), cci = T(tc.cci), c11 = T(tc.c11), cI = T(tc.cI), cIi = T(tc.cIi);
  T Cif[9];  // full C^-1
#pragma unroll
  for (int A = 0; A < 3; ++A)
#pragma unroll
    for (int B = 0; B < 3; ++B) Cif[A * 3 + B] = s.Ci[vidx(A, B)];
#pragma unroll
  for (int I = 0; I < 6; ++I) {
#pragma unroll
    for (int J = 0; J < 6; ++J) {
      const int A = vA(I), B = vB(I), Cc = vA(J), Dd = vB(J);
      const T dI = (I < 3) ? T(1) : T(0), dJ = (J < 3) ? T(1) : T(0);
      T v = cii * s.Ci[I] * s.Ci[J] + c1i * (dI * s.Ci[J] + s.Ci[I] * dJ) +
            cci * (s.C[I] * s.Ci[J] + s.Ci[I] * s.C[J]) + c11 * dI * dJ;
      if (I == J) v += cI * (I < 3 ? T(1) : T(0.5));
      v += cIi * T(0.5) * (Cif[A * 3 + Cc] * Cif[B * 3 + Dd] + Cif[A * 3 + Dd] * Cif[B * 3 + Cc]);
      D[I][J] = v;
    }
  }
}
// First elasticity tensor times w, AA_{jB,kD} = w (sum_{A,C} F_jA CC_ABCD F_kC + delta_jk S_BD),
// from the structure of CC above (no 6 x 6 intermediate): with G = F Ci, H = F C,
// b = F F^T and Q = G F^T,
//   F CC F^T = G (x) X + Y (x) G + c11 F (x) F + cI/2 (b_jk delta_BD + F_jD F_kB)
//              + cIi/2 (Q_jk Ci_BD + G_jD G_kB),
// X = cii G + c1i F + cci H, Y = c1i F + cci H ((P (x) R)_{jB,kD} = P_jB R_kD).
// Written as the packed upper triangle of the symmetric 9 x 9 matrix (index 3j+B).
template <typename T, typename F9>
__device__ __forceinline__ void first_elasticity_packed(const GPState<double> &s, const TangentCoeffs &tc,
                                                        const double S[6], double w, F9 &&put) {
  double Ci[9], Cf[9];
#pragma unroll
  for (int A = 0; A < 3; ++A)
#pragma unroll
    for (int B = 0; B < 3; ++B) {
      Ci[A * 3 + B] = s.Ci[vidx(A, B)];
      Cf[A * 3 + B] = s.C[vidx(A, B)];
    }
  double G[9], X[9], Y[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int B = 0; B < 3; ++B) {
      double g = 0.0, h = 0.0;
#pragma unroll
      for (int A = 0; A < 3; ++A) {
        g += s.F[j * 3 + A] * Ci[A * 3 + B];
        h += s.F[j * 3 + A] * Cf[A * 3 + B];
      }
      G[j * 3 + B] = g;
      Y[j * 3 + B] = w * (tc.c1i * s.F[j * 3 + B] + tc.cci * h);
      X[j * 3 + B] = w * tc.cii * g + Y[j * 3 + B];
    }
  double b[9], Q[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double bb = 0.0, q = 0.0;
#pragma unroll
      for (int A = 0; A < 3; ++A) {
        bb += s.F[j * 3 + A] * s.F[k * 3 + A];
        q += G[j * 3 + A] * s.F[k * 3 + A];
      }
      b[j * 3 + k] = bb;
      Q[j * 3 + k] = q;
    }
  const double w11 = w * tc.c11, wI = 0.5 * w * tc.cI, wIi = 0.5 * w * tc.cIi;
#pragma unroll
  for (int I = 0; I < 9; ++I)
#pragma unroll
    for (int J = I; J < 9; ++J) {
      const int j = I / 3, B = I % 3, k = J / 3, D = J % 3;
      double v = G[j * 3 + B] * X[k * 3 + D] + Y[j * 3 + B] * G[k * 3 + D] +
                 w11 * s.F[j * 3 + B] * s.F[k * 3 + D] +
                 wI * ((B == D ? b[j * 3 + k] : 0.0) + s.F[j * 3 + D] * s.F[k * 3 + B]) +
                 wIi * (Q[j * 3 + k] * Ci[B * 3 + D] + G[j * 3 + D] * G[k * 3 + B]);
      if (j == k) v += w * S[vidx(B, D)];
      put(I * 9 - (I * (I - 1)) / 2 + (J - I), v);
    }
}
// Strain energy density (hyperelastic.py:25-33 generalised), float64.
__device__ __forceinline__ double energy_density(double I1, double I2, double J, double Jm1,
                                                 const MatParams &m) {
  const double Jm23 = pow(J, -2.0 / 3.0);
  return m.w1 * (I1 * Jm23 - 3.0) + m.w2 * (I2 * Jm23 * Jm23 - 3.0) + 0.5 * m.kv * Jm1 * Jm1;
}

// B matrix of node a (Voigt rows, displacement component columns):
//   dE_AB/du_aj = (F_jA dN_B + F_jB dN_A)/2, shear rows engineering (x2).
template <typename T>
__device__ __forceinline__ void bmat(const T F[9], const T dN[3], T B[6][3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const T f0 = F[j * 3 + 0], f1 = F[j * 3 + 1], f2 = F[j * 3 + 2];
    B[0][j] = f0 * dN[0];
    B[1][j] = f1 * dN[1];
    B[2][j] = f2 * dN[2];
    B[3][j] = f0 * dN[1] + f1 * dN[0];
    B[4][j] = f1 * dN[2] + f2 * dN[1];
    B[5][j] = f0 * dN[2] + f2 * dN[0];
  }
}

// Gauss-point geometry: dN/dX [nen][3] and w*detJ from element coordinates.
// Sums start from -0.0 (x + -0.0 == x, so the compiler drops the first add) and skip the
// structural zeros of the reference derivatives.  C0 = 1: node 0's coordinates (x) or
// displacements (u) are exactly zero (element-relative values) and are skipped too.  For
// finite inputs the results equal the full sums (up to the sign of an exact zero).
template <typename T, int ET, int C0 = 0>
__device__ __forceinline__ T gp_geometry(int g, const T x[][3], T dNdX[][3]) {
  #pragma clang fp contract(on)
  constexpr int NEN = Elem<ET>::nen;
  double dr[3][NEN];
  Elem<ET>::dshape(g, dr);
  T Jm[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      T acc = T(-0.0);
#pragma unroll
      for (int c = C0; c < NEN; ++c)
        if (!Elem<ET>::dzero(m, c)) acc += x[c][i] * T(dr[m][c]);
      Jm[i * 3 + m] = acc;   // J_im = dx_i/dxi_m  (C3base.py:199-202)
    }
  const T dj = det3(Jm);
  T iJ[9];
  inv3(Jm, dj, iJ);        // iJ[m*3+i] = (J^-1)_mi
#pragma unroll
  for (int c = 0; c < NEN; ++c)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T v = T(-0.0);
#pragma unroll
      for (int m = 0; m < 3; ++m)
        if (!Elem<ET>::dzero(m, c)) v += iJ[m * 3 + i] * T(dr[m][c]);
      dNdX[c][i] = v;
    }
  return T(Elem<ET>::weight(g)) * dj;  // signed, C3base.py:214
}

template <typename T, int NEN, int C0 = 0>
__device__ __forceinline__ void disp_grad(const T u[][3], const T dNdX[][3], T H[9]) {
  #pragma clang fp contract(on)
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T acc = T(-0.0);
#pragma unroll
      for (int c = C0; c < NEN; ++c) acc += u[c][I] * dNdX[c][i];
      H[I * 3 + i] = acc;
    }
}

}  // namespace fea

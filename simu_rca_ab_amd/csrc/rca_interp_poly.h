// Generated by gen_interp_poly.py: polynomial forms of the Lanczos-type
// interpolation weights (see that file).  Pair p = 1..HI (taps p and 1-p):
// E_p(u) = kEvenLead[p-1] * (u^kDeg + sum_j kEven[p-1][j] u^j), O_p likewise.
#ifndef RCA_INTERP_POLY_H
#define RCA_INTERP_POLY_H

template <int KIND> struct InterpPoly;

// RCA_INTERP_LANCZOS_REF: sinc(x) sinc(x/2), taps -2..3, degree 4 in u; max |error| 1.72e-07 (FP32)
template <> struct InterpPoly<RCA_INTERP_LANCZOS_REF> {
  static constexpr int kDeg = 4;
  static constexpr float kEvenLead[3] = {-4.538312554e-02f, 9.044420719e-02f, -5.456493422e-02f};
  static constexpr float kEven[3][4] = {{-1.262934494e+01f, 5.704586983e+00f, 3.747802734e+00f, -3.344351530e+00f}, {-7.041288018e-01f, 1.929918289e+00f, 4.343402386e+00f, -3.438032627e+00f}, {4.201665223e-01f, -2.918062925e+00f, 5.831449986e+00f, -3.777424812e+00f}};
  static constexpr float kOddLead[3] = {5.264767632e-02f, -1.541781798e-02f, -1.791745983e-02f};
  static constexpr float kOdd[3][4] = {{2.644599533e+01f, -3.403206635e+01f, 1.821740723e+01f, -5.522166729e+00f}, {1.199570560e+01f, -5.511255646e+01f, 3.033367538e+01f, -7.508980274e+00f}, {-9.862776995e-01f, 3.793717384e+00f, 1.262475491e+00f, -2.877591848e+00f}};
};

// RCA_INTERP_LANCZOS3: sinc(x) sinc(x/3), taps -2..3, degree 4 in u; max |error| 1.28e-07 (FP32)
template <> struct InterpPoly<RCA_INTERP_LANCZOS3> {
  static constexpr int kDeg = 4;
  static constexpr float kEvenLead[3] = {-8.750166744e-03f, 2.266611904e-02f, -2.733590081e-02f};
  static constexpr float kEven[3][4] = {{-6.947605896e+01f, 5.129499817e+01f, -7.275265694e+00f, -2.474259138e+00f}, {-5.960213661e+00f, 2.473363304e+01f, -2.931310415e+00f, -2.809197426e+00f}, {-8.895660043e-01f, 3.159858942e+00f, 2.356008530e+00f, -3.299529076e+00f}};
  static constexpr float kOddLead[3] = {2.363377437e-02f, -1.697875746e-02f, 7.148234174e-03f};
  static constexpr float kOdd[3][4] = {{5.623522568e+01f, -6.214099503e+01f, 2.742187119e+01f, -6.746968746e+00f}, {1.060893536e+01f, -4.873975372e+01f, 2.687414360e+01f, -6.882340908e+00f}, {8.891702652e+00f, -4.266609955e+01f, 3.028146744e+01f, -7.787238121e+00f}};
};

// RCA_INTERP_LANCZOS4: sinc(x) sinc(x/4), taps -3..4, degree 3 in u; max |error| 4.28e-07 (FP32)
template <> struct InterpPoly<RCA_INTERP_LANCZOS4> {
  static constexpr int kDeg = 3;
  static constexpr float kEvenLead[4] = {-9.372168221e-03f, 1.341162249e-02f, 5.431956146e-03f, -2.740459703e-02f};
  static constexpr float kEven[4][3] = {{-6.619418335e+01f, 5.478969955e+01f, -1.389310360e+01f}, {-1.240829086e+01f, 5.389864349e+01f, -1.731214523e+01f}, {1.102910519e+01f, -4.966910172e+01f, 2.196163177e+01f}, {4.620043635e-01f, -1.809994578e+00f, -4.018648267e-01f}};
  static constexpr float kOddLead[4] = {-1.105787531e-01f, 8.962266892e-02f, -5.904270336e-02f, 3.037074767e-02f};
  static constexpr float kOdd[4][3] = {{-1.180344486e+01f, 1.219827938e+01f, -4.881210804e+00f}, {-1.871700883e+00f, 8.541777611e+00f, -4.469857693e+00f}, {-1.141825438e+00f, 5.499830723e+00f, -3.980075121e+00f}, {-1.028647542e+00f, 4.997512817e+00f, -3.781648874e+00f}};
};

#endif  // RCA_INTERP_POLY_H

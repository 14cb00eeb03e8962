// Generated by gen_interp_poly.py: polynomial forms of the Lanczos-type
// interpolation weights (see that file).  kEven[p-1][j], kOdd[p-1][j]: the
// coefficient of u^j of E_p and O_p; taps p = 1..HI and 1-p.
#ifndef RCA_INTERP_POLY_H
#define RCA_INTERP_POLY_H

constexpr int kPolyDeg = 4;

template <int KIND> struct InterpPoly;

// RCA_INTERP_LANCZOS_REF: sinc(x) sinc(x/2), taps -2..3; max |error| 1.35e-07 (FP32)
template <> struct InterpPoly<RCA_INTERP_LANCZOS_REF> {
  static constexpr float kEven[3][5] = {{5.731591582e-01f, -2.588919997e-01f, -1.700870097e-01f, 1.517771333e-01f, -4.538312554e-02f}, {-6.368436664e-02f, 1.745499223e-01f, 3.928355575e-01f, -3.109501302e-01f, 9.044420719e-02f}, {-2.292635851e-02f, 1.592239141e-01f, -3.181926906e-01f, 2.061149329e-01f, -5.456493422e-02f}};
  static constexpr float kOdd[3][5] = {{1.392320275e+00f, -1.791709185e+00f, 9.591041803e-01f, -2.907292545e-01f, 5.264767632e-02f}, {-1.849476099e-01f, 8.497153521e-01f, -4.676790833e-01f, 1.157720909e-01f, -1.541781798e-02f}, {1.767159067e-02f, -6.797377765e-02f, -2.262035385e-02f, 5.155913532e-02f, -1.791745983e-02f}};
};

// RCA_INTERP_LANCZOS3: sinc(x) sinc(x/3), taps -2..3; max |error| 1.34e-07 (FP32)
template <> struct InterpPoly<RCA_INTERP_LANCZOS3> {
  static constexpr float kEven[3][5] = {{6.079270840e-01f, -4.488397837e-01f, 6.365978718e-02f, 2.165018022e-02f, -8.750166744e-03f}, {-1.350949109e-01f, 5.606154799e-01f, -6.644143164e-02f, -6.367360801e-02f, 2.266611904e-02f}, {2.431708761e-02f, -8.637759089e-02f, -6.440361589e-02f, 9.019560367e-02f, -2.733590081e-02f}};
  static constexpr float kOdd[3][5] = {{1.329050541e+00f, -1.468626142e+00f, 6.480823159e-01f, -1.594563276e-01f, 2.363377437e-02f}, {-1.801265329e-01f, 8.275404572e-01f, -4.562895596e-01f, 1.168535948e-01f, -1.697875746e-02f}, {6.355997175e-02f, -3.049872816e-01f, 2.164590210e-01f, -5.566500127e-02f, 7.148234174e-03f}};
};

// RCA_INTERP_LANCZOS4: sinc(x) sinc(x/4), taps -3..4; max |error| 1.33e-07 (FP32)
template <> struct InterpPoly<RCA_INTERP_LANCZOS4> {
  static constexpr float kEven[4][5] = {{6.203830242e-01f, -5.134930015e-01f, 1.300904155e-01f, -8.578132838e-03f, -1.634843531e-03f}, {-1.664152294e-01f, 7.228511572e-01f, -2.318032086e-01f, 1.085147168e-02f, 5.271105561e-03f}, {5.990948528e-02f, -2.697711587e-01f, 1.186442822e-01f, 9.804795496e-03f, -9.003257379e-03f}, {-1.266087964e-02f, 4.956585541e-02f, 1.182122901e-02f, -3.283949196e-02f, 1.118992362e-02f}};
  static constexpr float kOdd[4][5] = {{1.305212140e+00f, -1.348976731e+00f, 5.413569212e-01f, -1.193343848e-01f, 1.584221236e-02f}, {-1.677483320e-01f, 7.656254768e-01f, -4.019347429e-01f, 9.692948312e-02f, -1.322076097e-02f}, {6.741748750e-02f, -3.247851431e-01f, 2.359021753e-01f, -6.401429325e-02f, 8.995461278e-03f}, {-3.124131449e-02f, 1.518085450e-01f, -1.153085232e-01f, 3.287364915e-02f, -4.528682679e-03f}};
};

#endif  // RCA_INTERP_POLY_H

// Generated by gen_j1_modphase.py: J1(x) = m(y) sqrt(y) cos(x - 3 pi/4 + 2 pi p(y)),
// y = 1/x, x >= 4; coefficients of y^0..y^7 (least squares against
// scipy.special.j1 / y1: m relative error 2.3e-09, p 5.9e-11 revolutions).
#pragma once

namespace pii {

constexpr int kJ1ModPhaseDeg = 7;
constexpr float kJ1Mod[8] = {7.978845590e-01f, 5.255630962e-07f, 1.495662627e-01f, 1.088739342e-03f, -1.702768300e-01f, 1.223219910e-01f, 2.355439359e-01f, -3.818697245e-01f};
constexpr float kJ1Phase[8] = {0.000000000e+00f, 5.968312414e-02f, -2.284266263e-06f, -2.601867251e-02f, -1.888771956e-03f, 8.075361991e-02f, -1.435187365e-01f, 9.609764418e-02f};

}  // namespace pii

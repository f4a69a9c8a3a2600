// madam_ns_coef.h -- coefficient schedule of the sign iteration used by the
// mixed-precision PSD projection (madam_ns.hip), shared by host and device.
//
// Step k maps x -> a_k x + b_k x^3 + c_k x^5 on the spectrum of the FP32
// iterate, which starts inside [-1, 1] (W / |W|_F) and must end at +-1 for
// every |x| >= 1e-6.  Two constraints shape the schedule besides speed:
//  * stability: each polynomial keeps values up to 1.02 times the top of
//    its input interval inside the next interval, so values rounded above
//    the top do not grow step after step;
//  * no squashing: a value never drops below ~0.39 of the running maximum
//    once it has reached it.  A minimax polynomial for a wide interval
//    [l, u] attains its lowest value l' (tiny) also at interior points, so
//    an eigenvalue already near 1 can be mapped to ~1e-4; FP32 rounding
//    (~1e-7 absolute) then mixes its eigenvector with others at relative
//    size 1e-3, and the final W_+ carries that error (measured 1e-4 |W|_F).
// Steps 0-8 are one polynomial, the fastest growth of small values (x3.62
// per step) under both constraints; steps 9-12 are minimax polynomials for
// the narrowing intervals [0.106, 1], [0.56, 1.44], [0.94, 1.06],
// [0.9998, 1.0002].  After the 13 steps every x in [1e-6, 1] is within
// 1e-6 of 1, and the largest intermediate value is 1.44.
#pragma once

#define MADAM_NS_STEPS 13
#define MADAM_NS_COEF_INIT                                                  \
  {{3.6190245f, -7.9645433f, 5.217582f},  {3.6190245f, -7.9645433f, 5.217582f},  \
   {3.6190245f, -7.9645433f, 5.217582f},  {3.6190245f, -7.9645433f, 5.217582f},  \
   {3.6190245f, -7.9645433f, 5.217582f},  {3.6190245f, -7.9645433f, 5.217582f},  \
   {3.6190245f, -7.9645433f, 5.217582f},  {3.6190245f, -7.9645433f, 5.217582f},  \
   {3.6190243f, -7.9645433f, 5.217582f},  {5.4176984f, -12.800723f, 8.602348f},  \
   {2.0927296f, -1.4410863f, 0.37352878f}, {1.8604435f, -1.21661f, 0.35627013f}, \
   {1.8562274f, -1.2127392f, 0.35651124f}}

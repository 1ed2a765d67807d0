// rh_opcount.hpp -- an op-counting scalar for the host build of the kernel
// math (rh_math.hpp, rh_mesh.hpp).  Instantiating the templated integrals
// with R = opc::OpC counts every arithmetic operation they perform in R as
// written, with the branches a given point takes: add, sub, mul, div
// (reciprocals included), sqrt, fused multiply-add, and each log and atan2 as
// one call.  Host only; the device never sees this type.
#pragma once

#include <cmath>

#include "rh_math.hpp"

namespace opc {

struct Counts {
    double add = 0, mul = 0, div = 0, sqrt = 0, fma = 0, log = 0, atan2 = 0;
};

// The counter the operators add to (null: not counting).
inline Counts*& counter()
{
    static thread_local Counts* c = nullptr;
    return c;
}

struct OpC {
    double v = 0.0;
    OpC() = default;
    OpC(double x) : v(x) {}  // NOLINT: implicit, as R(x) in the templates
    explicit operator double() const { return v; }
};

#define OPC_BIN(op, field)                                              \
    inline OpC operator op(OpC a, OpC b)                                \
    {                                                                   \
        if (counter()) counter()->field += 1;                           \
        return OpC(a.v op b.v);                                         \
    }
OPC_BIN(+, add)
OPC_BIN(-, add)
OPC_BIN(*, mul)
OPC_BIN(/, div)
#undef OPC_BIN
inline OpC operator-(OpC a) { return OpC(-a.v); }
inline OpC& operator+=(OpC& a, OpC b) { return a = a + b; }
inline OpC& operator-=(OpC& a, OpC b) { return a = a - b; }
inline OpC& operator*=(OpC& a, OpC b) { return a = a * b; }
inline bool operator<(OpC a, OpC b) { return a.v < b.v; }
inline bool operator>(OpC a, OpC b) { return a.v > b.v; }
inline bool operator<=(OpC a, OpC b) { return a.v <= b.v; }
inline bool operator>=(OpC a, OpC b) { return a.v >= b.v; }

inline OpC fabs(OpC a) { return OpC(std::fabs(a.v)); }
inline OpC fma(OpC a, OpC b, OpC c)
{
    if (counter()) counter()->fma += 1;
    return OpC(std::fma(a.v, b.v, c.v));
}
inline OpC rh_sqrt(OpC a)
{
    if (counter()) counter()->sqrt += 1;
    return OpC(std::sqrt(a.v));
}
inline OpC rh_rcp(OpC a)
{
    if (counter()) counter()->div += 1;
    return OpC(1.0 / a.v);
}
// ln(num / den): one log and one division
inline OpC rh_log_ratio(OpC num, OpC den)
{
    if (counter()) {
        counter()->log += 1;
        counter()->div += 1;
    }
    return OpC(std::log(num.v / den.v));
}
inline OpC rh_log1p(OpC a)
{
    if (counter()) counter()->log += 1;
    return OpC(std::log1p(a.v));
}
inline OpC rh_atan2(OpC y, OpC x)
{
    if (counter()) counter()->atan2 += 1;
    return OpC(std::atan2(y.v, x.v));
}
inline OpC rh_atan2_wave(OpC y, OpC x) { return rh_atan2(y, x); }

// The FP64 edge term of rh_math.hpp (edge_L for double), in the counting type:
// the same operations on the same branch.
inline OpC edge_L(const rh::V3<OpC>& ra, OpC Ra, const rh::V3<OpC>& rb, OpC Rb, OpC dab, OpC len)
{
    const OpC q = len * (len + (Ra + Rb));
    OpC den = fma(Ra, Rb, dab);
    OpC num = den + q;
    if (dab < OpC(0.0)) {
        const rh::V3<OpC> c = rh::cross(ra, rb);
        den = rh::dot(c, c);
        num = fma(q, fma(Ra, Rb, -dab), den);
    }
    return rh_log_ratio(num, den);
}

}  // namespace opc

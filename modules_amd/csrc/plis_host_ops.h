// plis_host_ops.h -- PLIS vector operations over host arrays, in the order of NLopt's luksan_mxudot__ /
// mxudir__ / mxvscl__ / mxvdif__ / mxuneg__ / mxvmax__ (potential/LuksanPLIS.cpp:106-370), no FP contraction.
#pragma once
#include <cmath>
#include <cstring>

#include "plis_core.h"

namespace ccgpu {

#pragma clang fp contract(off)
// Vector operations over host arrays in the order of luksan_mxudot__/mxudir__/... (LuksanPLIS.cpp:106-370).
struct HostPlisOps {
    static constexpr bool kDeferTwoLoop = false;
    int n, mf;
    double *x, *g, *s, *xo, *go;  // xo, go: mf x n
    double *uo, *vo;              // mf each
    double u(int i) const { return uo[i]; }
    double v(int i) const { return vo[i]; }
    void set_u(int i, double val) const { uo[i] = val; }
    void set_v(int i, double val) const { vo[i] = val; }
    void shift_u(int k) const
    {
        for (int i = k - 1; i >= 1; --i) uo[i] = uo[i - 1];
    }
    double two_loop(int k, int head, double b)
    {
        return PlisTwoLoopSequential(*this, k, head, mf, b);
    }
    double *vec(int id) const
    {
        if (id == kVecX) return x;
        if (id == kVecG) return g;
        if (id == kVecS) return s;
        id -= kVecXO;
        return id < mf ? xo + (size_t)id * n : go + (size_t)(id - mf) * n;
    }
    double dot(int a, int b) const
    {
        const double *pa = vec(a), *pb = vec(b);
        double t = 0.;
        for (int i = 0; i < n; i++) t += pa[i] * pb[i];
        return t;
    }
    void axpy(int z, double a, int xv, int yv) const  // z = y + a x
    {
        double *pz = vec(z);
        const double *px = vec(xv), *py = vec(yv);
        for (int i = 0; i < n; i++) pz[i] = py[i] + a * px[i];
    }
    void neg(int src, int dst) const
    {
        const double *ps = vec(src);
        double *pd = vec(dst);
        for (int i = 0; i < n; i++) pd[i] = -ps[i];
    }
    void copy(int src, int dst) const { std::memcpy(vec(dst), vec(src), sizeof(double) * n); }
    void scal(double a, int xv) const
    {
        double *p = vec(xv);
        for (int i = 0; i < n; i++) p[i] = a * p[i];
    }
    void diff(int xv, int yv, int z) const  // z = x - y
    {
        const double *px = vec(xv), *py = vec(yv);
        double *pz = vec(z);
        for (int i = 0; i < n; i++) pz[i] = px[i] - py[i];
    }
    // the new difference pair in slots (sv, yv): s = x - s, y = g - y
    void form_pair(int sv, int yv) const
    {
        diff(kVecX, sv, sv);
        diff(kVecG, yv, yv);
    }
    // the start point of a line search into slots (sv, yv), and with a trial x = s + r S
    void save_start(int sv, int yv, double r, bool trial) const
    {
        copy(kVecX, sv);
        copy(kVecG, yv);
        if (trial) axpy(kVecX, r, kVecS, sv);
    }
    double maxabs(int xv) const
    {
        const double *p = vec(xv);
        double m = 0.;
        for (int i = 0; i < n; i++) m = m > fabs(p[i]) ? m : fabs(p[i]);
        return m;
    }
    double l1(int xv) const
    {
        const double *p = vec(xv);
        double t = 0;
        for (int i = 0; i < n; i++) t += fabs(p[i]);
        return t;
    }
};

}  // namespace ccgpu

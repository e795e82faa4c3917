/*
 * Reverse-mode recording scalar for the gradient oracle (wrfdyn_gradient.cpp).
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/.
 *
 * `ad` carries a double value and the index of the tape node that produced
 * it; index 0 marks a constant.  While recording is on, every operation with
 * a non-constant operand appends one node holding the indices of its (at most
 * two) operands and the partial derivatives with respect to them, evaluated
 * at the operand values.  A reverse sweep over the tape then accumulates
 * adjoints from the outputs back to the leaves: that is the exact derivative
 * of the recorded double-precision expressions.  Values are computed with the
 * same IEEE operations, in the same order, as the double build, so a recorded
 * run gives the same bits as the plain oracle.
 *
 * Derivative conventions at the points where the recorded functions are not
 * differentiable follow jax's:
 *   |x|            -> sign(x), 0 at x = 0
 *   comparisons / selects -> the derivative of the selected operand
 *   sign(x)        -> 0 (the caller builds it as a constant)
 *
 * Single-threaded: the tape is one global object.
 */
#ifndef WRFDYN_AD_H
#define WRFDYN_AD_H

#include <math.h>

#include <vector>

#include "wrfdyn_math.h"

struct ad_node {
    int a, b;
    double da, db;
};

struct ad_tape {
    std::vector<ad_node> nodes;
    bool recording = false;
    void reset() {
        nodes.clear();
        nodes.push_back({0, 0, 0.0, 0.0}); /* node 0: constants */
    }
};

extern ad_tape g_ad_tape;

struct ad {
    double v;
    int id;
    ad() : v(0.0), id(0) {}
    ad(double x) : v(x), id(0) {}
    ad(double x, int n) : v(x), id(n) {}
};

static inline int ad_push(int a, double da, int b, double db) {
    if (!g_ad_tape.recording || (a == 0 && b == 0)) return 0;
    g_ad_tape.nodes.push_back({a, b, da, db});
    return (int)g_ad_tape.nodes.size() - 1;
}

/* A new independent variable. */
static inline ad ad_leaf(double x) {
    g_ad_tape.nodes.push_back({0, 0, 0.0, 0.0});
    return ad(x, (int)g_ad_tape.nodes.size() - 1);
}

static inline ad operator+(ad x, ad y) { return ad(x.v + y.v, ad_push(x.id, 1.0, y.id, 1.0)); }
static inline ad operator+(ad x, double y) { return ad(x.v + y, ad_push(x.id, 1.0, 0, 0.0)); }
static inline ad operator+(double x, ad y) { return ad(x + y.v, ad_push(y.id, 1.0, 0, 0.0)); }
static inline ad operator-(ad x, ad y) { return ad(x.v - y.v, ad_push(x.id, 1.0, y.id, -1.0)); }
static inline ad operator-(ad x, double y) { return ad(x.v - y, ad_push(x.id, 1.0, 0, 0.0)); }
static inline ad operator-(double x, ad y) { return ad(x - y.v, ad_push(y.id, -1.0, 0, 0.0)); }
static inline ad operator-(ad x) { return ad(-x.v, ad_push(x.id, -1.0, 0, 0.0)); }
static inline ad operator*(ad x, ad y) { return ad(x.v * y.v, ad_push(x.id, y.v, y.id, x.v)); }
static inline ad operator*(ad x, double y) { return ad(x.v * y, ad_push(x.id, y, 0, 0.0)); }
static inline ad operator*(double x, ad y) { return ad(x * y.v, ad_push(y.id, x, 0, 0.0)); }
static inline ad operator/(ad x, ad y) {
    double q = x.v / y.v;
    return ad(q, ad_push(x.id, 1.0 / y.v, y.id, -q / y.v));
}
static inline ad operator/(ad x, double y) { return ad(x.v / y, ad_push(x.id, 1.0 / y, 0, 0.0)); }
static inline ad operator/(double x, ad y) {
    double q = x / y.v;
    return ad(q, ad_push(y.id, -q / y.v, 0, 0.0));
}
static inline ad &operator+=(ad &x, ad y) { return x = x + y; }
static inline ad &operator-=(ad &x, ad y) { return x = x - y; }
static inline ad &operator*=(ad &x, ad y) { return x = x * y; }
static inline ad &operator/=(ad &x, ad y) { return x = x / y; }

#define AD_CMP(op)                                                               \
    static inline bool operator op(ad x, ad y) { return x.v op y.v; }            \
    static inline bool operator op(ad x, double y) { return x.v op y; }          \
    static inline bool operator op(double x, ad y) { return x op y.v; }
AD_CMP(<)
AD_CMP(>)
AD_CMP(<=)
AD_CMP(>=)
AD_CMP(==)
AD_CMP(!=)
#undef AD_CMP

static inline ad fabs(ad x) {
    double s = x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : 0.0);
    return ad(fabs(x.v), ad_push(x.id, s, 0, 0.0));
}

/* d/dx x^y = y x^y / x (x > 0 on the recorded path: the equation of state
 * clamps its arguments positive first). */
static inline ad wd_pow(ad x, double y) {
    double p = wd_pow(x.v, y);
    return ad(p, ad_push(x.id, (y * p) / x.v, 0, 0.0));
}

static inline ad wd_sin(ad x) { return ad(wd_sin(x.v), ad_push(x.id, cos(x.v), 0, 0.0)); }
static inline ad wd_log(ad x) { return ad(wd_log(x.v), ad_push(x.id, 1.0 / x.v, 0, 0.0)); }
static inline ad sqrt(ad x) {
    const double r = sqrt(x.v);
    return ad(r, ad_push(x.id, 0.5 / r, 0, 0.0));
}

/* Adjoints of every node, seeded by the caller, swept from the last node back. */
static inline void ad_reverse(std::vector<double> &adj) {
    const std::vector<ad_node> &t = g_ad_tape.nodes;
    for (size_t n = t.size() - 1; n >= 1; --n) {
        double g = adj[n];
        if (g == 0.0) continue;
        const ad_node &e = t[n];
        if (e.a) adj[e.a] += e.da * g;
        if (e.b) adj[e.b] += e.db * g;
    }
}

#endif

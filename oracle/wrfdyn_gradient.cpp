/*
 * Gradient oracle: the vector-Jacobian product of the CPU oracle's large
 * step, by reverse-mode differentiation of wrfdyn_oracle.c itself.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for adjoint kernels, never linked
 * into or called by the product.
 *
 * The reference has no gradient of its own (no jax.grad/vjp anywhere in its
 * source); what its step would give under jax.grad is the derivative of the
 * same FP64 expressions that wrfdyn_oracle.c restates
 * (src/gpuwrf/runtime/operational_mode.py:3319-3534).  This file compiles
 * wrfdyn_oracle.c unchanged as C++ with `real` the recording scalar of
 * wrfdyn_ad.h, so the gradient is that of exactly the oracle's expressions,
 * with jax's conventions at the non-differentiable points (sign has a zero
 * derivative, |x| has derivative sign(x), selects pass the chosen operand's
 * derivative).  It is pinned by central finite differences of the double
 * oracle (tests/test_oracle_gradient.py).
 *
 * Every one of the eleven state arrays is an independent input, as the
 * reference's state dict holds them (p_total and p_pert, ph_total and
 * ph_pert, mu_total and mu_pert separately, and mudf).
 *
 * Strategy: checkpoint the state at the start of every large step (a plain,
 * unrecorded forward run), then for steps n-1 .. 0 record one large step from
 * its checkpoint, seed the adjoints of its outputs and sweep back to its
 * inputs.  Single-threaded.
 */
#include "wrfdyn_ad.h"

#include <stdlib.h>
#include <string.h>

#include "wrfdyn_math.h"

ad_tape g_ad_tape;

#define ORC_AD 1
typedef ad real;
/* jnp.sign as a constant: its derivative is zero everywhere. */
static inline real sgn(real x) { return real(x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : x.v)); }

#include "wrfdyn_oracle.c"

namespace {

constexpr int kFields = 11;

struct sizes {
    size_t n[kFields];
    explicit sizes(const orc_grid *g) {
        const size_t nx = g->nx, ny = g->ny, nz = g->nz;
        const size_t M = nz * ny * nx, W = (nz + 1) * ny * nx, C = ny * nx;
        const size_t U = nz * ny * (nx + 1), V = nz * (ny + 1) * nx;
        const size_t all[kFields] = {U, V, W, M, M, M, W, W, C, C, C};
        for (int f = 0; f < kFields; ++f) n[f] = all[f];
    }
};

/* The fields of a state struct in a fixed order (u, v, w, theta, p_total,
 * p_pert, ph_total, ph_pert, mu_total, mu_pert, mudf). */
template <class S, class T>
void fields(S *s, T **out) {
    T *all[kFields] = {s->u, s->v, s->w, s->theta, s->p_total, s->p_pert,
                       s->ph_total, s->ph_pert, s->mu_total, s->mu_pert, s->mudf};
    for (int f = 0; f < kFields; ++f) out[f] = all[f];
}

struct ad_state {
    std::vector<ad> a[kFields];
    rstate rs;
    explicit ad_state(const sizes &z) {
        for (int f = 0; f < kFields; ++f) a[f].resize(z.n[f]);
        rs = {a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(),
              a[6].data(), a[7].data(), a[8].data(), a[9].data(), a[10].data()};
    }
};

/* One large step of the oracle on `st` (values only or recorded). */
int step(const orc_grid *g, const orc_namelist *nl, ad_state &st) {
    ws_t W;
    if (ws_alloc(&W, g->nx, g->ny, g->nz) != 0) {
        ws_free(&W);
        return -1;
    }
    large_step(g, nl, &st.rs, &W);
    ws_free(&W);
    return 0;
}

}  // namespace

extern "C" {

/* Vector-Jacobian product of n_steps large steps.
 *   in       state at the start (double, not modified)
 *   bar_out  adjoint of the state after n_steps (not modified)
 *   bar_in   receives the adjoint of the starting state
 *   out      receives the state after n_steps (may be NULL)
 * Returns 0, or -1 on allocation failure, -2/-3 on an unsupported grid or
 * namelist (as orc_run). */
int orc_vjp(const orc_grid *g, const orc_namelist *nl, const orc_state *in, int n_steps,
            const orc_state *bar_out, orc_state *bar_in, orc_state *out) {
    if (g->nx < 4 || g->ny < 1 || g->nz < 4) return -2;
    if (!nl->top_lid) return -3;
    if (n_steps < 0) return -2;
    const sizes z(g);
    double *pin[kFields], *pbo[kFields], *pbi[kFields];
    fields(const_cast<orc_state *>(in), pin);
    fields(const_cast<orc_state *>(bar_out), pbo);
    fields(bar_in, pbi);

    /* checkpoints: ck[s] is the state entering step s */
    std::vector<std::vector<double>> ck((size_t)(n_steps + 1) * kFields);
    for (int f = 0; f < kFields; ++f) ck[f].assign(pin[f], pin[f] + z.n[f]);
    g_ad_tape.recording = false;
    for (int s = 0; s < n_steps; ++s) {
        ad_state st(z);
        for (int f = 0; f < kFields; ++f)
            for (size_t q = 0; q < z.n[f]; ++q) st.a[f][q] = ad(ck[(size_t)s * kFields + f][q]);
        if (step(g, nl, st) != 0) return -1;
        for (int f = 0; f < kFields; ++f) {
            std::vector<double> &c = ck[(size_t)(s + 1) * kFields + f];
            c.resize(z.n[f]);
            for (size_t q = 0; q < z.n[f]; ++q) c[q] = st.a[f][q].v;
        }
    }
    if (out) {
        double *po[kFields];
        fields(out, po);
        for (int f = 0; f < kFields; ++f)
            memcpy(po[f], ck[(size_t)n_steps * kFields + f].data(), z.n[f] * sizeof(double));
    }

    std::vector<double> bar[kFields];
    for (int f = 0; f < kFields; ++f) bar[f].assign(pbo[f], pbo[f] + z.n[f]);
    std::vector<double> adj;
    for (int s = n_steps - 1; s >= 0; --s) {
        g_ad_tape.reset();
        g_ad_tape.recording = true;
        ad_state st(z);
        std::vector<int> leaf[kFields];
        for (int f = 0; f < kFields; ++f) {
            leaf[f].resize(z.n[f]);
            for (size_t q = 0; q < z.n[f]; ++q) {
                st.a[f][q] = ad_leaf(ck[(size_t)s * kFields + f][q]);
                leaf[f][q] = st.a[f][q].id;
            }
        }
        int rc = step(g, nl, st);
        g_ad_tape.recording = false;
        if (rc != 0) return -1;
        adj.assign(g_ad_tape.nodes.size(), 0.0);
        for (int f = 0; f < kFields; ++f)
            for (size_t q = 0; q < z.n[f]; ++q)
                if (st.a[f][q].id) adj[st.a[f][q].id] += bar[f][q];
        ad_reverse(adj);
        for (int f = 0; f < kFields; ++f)
            for (size_t q = 0; q < z.n[f]; ++q) bar[f][q] = adj[leaf[f][q]];
    }
    g_ad_tape.nodes.clear();
    g_ad_tape.nodes.shrink_to_fit();
    for (int f = 0; f < kFields; ++f) memcpy(pbi[f], bar[f].data(), z.n[f] * sizeof(double));
    return 0;
}

/* Vector-Jacobian product of one piece of a large step (ORC_PIECE_*).
 *   arrays  the orc_array_count() arrays before the piece (not modified)
 *   bar     in: the adjoint of every array after the piece;
 *           out: the adjoint of every array before it
 * Returns 0 or a negative error. */
int orc_piece_vjp(const orc_grid *g, const orc_namelist *nl, int piece, int rk_step, double dts,
                  double *const *arrays, double *const *bar) {
    if (g->nx < 4 || g->ny < 1 || g->nz < 4) return -2;
    std::vector<std::vector<ad>> a(N_ARRAYS);
    std::vector<std::vector<int>> leaf(N_ARRAYS);
    std::vector<real *> ptr(N_ARRAYS);
    g_ad_tape.reset();
    g_ad_tape.recording = true;
    for (int k = 0; k < N_ARRAYS; ++k) {
        const size_t n = extent_len(g->nx, g->ny, g->nz, array_extent[k]);
        a[k].resize(n);
        leaf[k].resize(n);
        for (size_t q = 0; q < n; ++q) {
            a[k][q] = ad_leaf(arrays[k][q]);
            leaf[k][q] = a[k][q].id;
        }
        ptr[k] = a[k].data();
    }
    rstate S;
    ws_t W;
    bind_arrays(ptr.data(), &S, &W);
    int rc = run_piece(g, nl, piece, rk_step, dts, &S, &W);
    g_ad_tape.recording = false;
    if (rc != 0) return rc;
    std::vector<double> adj(g_ad_tape.nodes.size(), 0.0);
    for (int k = 0; k < N_ARRAYS; ++k)
        for (size_t q = 0; q < a[k].size(); ++q)
            if (a[k][q].id) adj[a[k][q].id] += bar[k][q];
    ad_reverse(adj);
    for (int k = 0; k < N_ARRAYS; ++k)
        for (size_t q = 0; q < a[k].size(); ++q) bar[k][q] = adj[leaf[k][q]];
    g_ad_tape.nodes.clear();
    g_ad_tape.nodes.shrink_to_fit();
    return 0;
}

}  // extern "C"

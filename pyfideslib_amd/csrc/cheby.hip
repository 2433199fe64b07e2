// Chebyshev series sum_k c_k T_k(y), y = (2x - (a + b)) / (b - a), by
// Paterson-Stockmeyer in the Chebyshev basis -- the polynomial step of
// bootstrapping's EvalMod (FIDESlib evalChebyshevSeries,
// src/CKKS/ApproxModEval.cu:370-700, OpenFHE EvalChebyshevSeriesPS), built
// from the engine's ops (FIXEDMANUAL):
//   baby steps  T_1 .. T_m (m a power of two near sqrt(deg)):
//               T_2i = 2 T_i^2 - 1,  T_2i+1 = 2 T_i T_i+1 - T_1
//   giant steps G_t = T_(m 2^t) by doubling, G_t = 2 G_(t-1)^2 - 1
//   division    p = q T_M + r with M = m 2^t:  c_M T_M = T_M (c_M T_0) and
//               c_(M+j) T_(M+j) = T_M (2 c_(M+j) T_j) - c_(M+j) T_(M-j)
//   base        sum_(j<m) c_j T_j with scalar products at noise degree 2,
//               one rescale, then the constant.
// Scales are exact: FIXEDMANUAL rescales divide by primes that are only near
// 2^scale_bits, and the repeated squarings of the giant steps double that
// drift each time, so terms meeting in an addition are first put on one
// tracked scale.  A base sum steers every product c_j T_j onto a requested
// scale (integer factors round(c_j q_l S / S_j)); a division node asks its
// quotient for the scale that makes q T_M land on the node's scale, and its
// remainder for that same scale; T_(2i+1) = 2 T_i T_(i+1) - T_1 steers T_1
// onto the product's scale through one of the levels it drops anyway; the
// constants are added at the tracked scale.  A structure-only pass gives the
// level of every subtree so the quotient's scale can be chosen before it is
// computed.  The series lands on ScalingFactorReal of its output level.
// Operands of an add or a product are brought to the lower level by dropping
// limbs (FIXEDMANUAL keeps the scale near 2^scale_bits at every level).
// The operation order is this file's own, so results are checked against the
// FP64 value of the series, not residue by residue against the reference.
#include <cmath>

#include "context.hpp"

namespace fides {

namespace {

using CtPtr = std::unique_ptr<Ciphertext>;

struct Cheb {
    Context& c;
    std::vector<CtPtr> T;  // T[1..m]
    std::vector<CtPtr> G;  // G[t] = T_(m 2^t)
    int m = 1;

    CtPtr fresh() { return ct_create(c); }
    CtPtr copy(const Ciphertext& a) {
        CtPtr o = fresh();
        ct_copy(c, a, *o);
        return o;
    }
    // lower-level copies of a and b (level of the lower one)
    void align(CtPtr& a, CtPtr& b) {
        if (a->level > b->level) drop_to_level(c, *a, b->level);
        if (b->level > a->level) drop_to_level(c, *b, a->level);
    }
    // 2 a b - s (s steered onto the product's scale) or 2 a b - k
    CtPtr two_ab_minus(const Ciphertext& a, const Ciphertext& b, const Ciphertext* s, double k) {
        CtPtr p = fresh();
        eval_mult(c, a, b, *p, true);
        eval_add(c, *p, *p, *p);
        if (s) {
            if (s->level <= p->level) throw Error(E_STATE, "Chebyshev: subtrahend has no level to spare");
            CtPtr u = fresh();
            mult_const_to_scale(c, *s, p->level + 1, 1.0, p->noise, *u);
            rescale(c, *u);
            eval_sub(c, *p, *u, *p);
        } else {
            add_const_at_scale(c, *p, -k, *p);
        }
        return p;
    }
    static void trim(std::vector<double>& cf) {
        while (!cf.empty() && cf.back() == 0.0) cf.pop_back();
    }
    // level of base(cf): the lowest level among the T_j it uses, less the rescale
    int base_level(const std::vector<double>& cf) const {
        int lo = T[1]->level;
        for (size_t j = 1; j < cf.size(); ++j)
            if (cf[j] != 0.0) lo = std::min(lo, T[j]->level);
        return lo - 1;
    }
    // p = q T_M + r in the Chebyshev basis (cf.size() <= 2M)
    static void divide(const std::vector<double>& cf, int M, std::vector<double>& q, std::vector<double>& r) {
        q.assign(cf.size() - M, 0.0);
        r.assign(cf.begin(), cf.begin() + M);
        q[0] = cf[M];
        for (size_t j = 1; j < q.size(); ++j) {
            q[j] = 2.0 * cf[M + j];
            r[M - j] -= cf[M + j];
        }
    }
    // level of rec(cf, t), from the structure alone
    int rec_level(std::vector<double> cf, int t) const {
        trim(cf);
        if ((int)cf.size() <= m + 1 || t < 0) return base_level(cf);
        const int M = m << t;
        if ((int)cf.size() <= M) return rec_level(cf, t - 1);
        std::vector<double> q, r;
        divide(cf, M, q, r);
        const int pl = std::min(rec_level(q, t - 1), G[t]->level) - 1;
        return std::min(pl, rec_level(r, t - 1));
    }
    // sum_(j<n) cf[j] T_j at scale S (n <= m + 1; T_m taken from the baby steps)
    CtPtr base(const std::vector<double>& cf, double S) {
        const int lo = base_level(cf) + 1;
        std::vector<const Ciphertext*> xs;
        std::vector<double> vs;
        for (size_t j = 1; j < cf.size(); ++j)
            if (cf[j] != 0.0) {
                xs.push_back(T[j].get());
                vs.push_back(cf[j]);
            }
        CtPtr acc = fresh();
        if (xs.size() > 1)  // one launch for the whole sum
            lincomb_to_scale(c, xs, vs, lo, S, *acc);
        else if (xs.size() == 1)
            mult_const_to_scale(c, *xs[0], lo, vs[0], S, *acc);
        else  // constant: T_1 * 0 + c_0, as the reference's degree-0 case
            mult_const_to_scale(c, *T[1], lo, 0.0, S, *acc);
        rescale(c, *acc);
        if (!cf.empty() && cf[0] != 0.0) add_const_at_scale(c, *acc, cf[0], *acc);
        return acc;
    }
    // sum_j cf[j] T_j at scale S, for cf.size() <= m 2^(t+1)
    CtPtr rec(std::vector<double> cf, int t, double S) {
        trim(cf);
        if ((int)cf.size() <= m + 1 || t < 0) return base(cf, S);
        const int M = m << t;
        if ((int)cf.size() <= M) return rec(cf, t - 1, S);
        std::vector<double> q, r;
        divide(cf, M, q, r);
        // the quotient's scale makes q T_M land on S after the product's rescale
        const int lvl = std::min(rec_level(q, t - 1), G[t]->level);
        CtPtr Q = rec(q, t - 1, S * c.mod_reduce_factor[lvl] / G[t]->noise);
        CtPtr p = fresh();
        eval_mult(c, *Q, *G[t], *p, true);
        CtPtr R = rec(r, t - 1, p->noise);
        align(p, R);
        eval_add(c, *p, *R, *p);
        return p;
    }
};

}  // namespace

void eval_chebyshev_series(Context& c, const Ciphertext& x_in, const std::vector<double>& coeffs, double a, double b,
                           Ciphertext& out) {
    if (x_in.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (c.scaling_technique != FIXEDMANUAL) throw Error(E_STATE, "Chebyshev series runs under FIXEDMANUAL");
    if (coeffs.empty()) throw Error(E_ARG, "no coefficients");
    if (!(b > a)) throw Error(E_ARG, "need a < b");
    int deg = (int)coeffs.size() - 1;
    while (deg > 0 && coeffs[deg] == 0.0) --deg;
    Cheb S{c, {}, {}, 1};
    // y = (2x - (a + b)) / (b - a); no level spent when [a, b] = [-1, 1]
    CtPtr y = S.copy(x_in);
    if (y->noise_level == 2) rescale(c, *y);
    if (std::fabs(a + 1.0) > 1e-12 || std::fabs(b - 1.0) > 1e-12) {
        CtPtr t = S.fresh();
        mult_const_to_scale(c, *y, y->level, 2.0 / (b - a), y->noise, *t);
        rescale(c, *t);
        add_const_at_scale(c, *t, -(a + b) / (b - a), *t);
        y = std::move(t);
    }
    // baby steps
    while ((S.m + 1) * (S.m + 1) <= deg + 1 && S.m < 64) S.m <<= 1;
    S.T.resize(S.m + 1);
    S.T[1] = std::move(y);
    for (int i = 2; i <= S.m; ++i) {
        if (i % 2 == 0)
            S.T[i] = S.two_ab_minus(*S.T[i / 2], *S.T[i / 2], nullptr, 1.0);
        else
            S.T[i] = S.two_ab_minus(*S.T[i / 2], *S.T[i / 2 + 1], S.T[1].get(), 0.0);
    }
    // giant steps T_(m 2^t) while m 2^t <= deg
    int tmax = -1;
    S.G.push_back(S.copy(*S.T[S.m]));
    if (S.m <= deg) tmax = 0;
    while ((S.m << (tmax + 1)) <= deg) {
        S.G.push_back(S.two_ab_minus(*S.G.back(), *S.G.back(), nullptr, 1.0));
        ++tmax;
    }
    std::vector<double> cf(coeffs.begin(), coeffs.begin() + deg + 1);
    const int lout = S.rec_level(cf, tmax);
    if (lout < 0) throw Error(E_LEVEL, "not enough levels for the Chebyshev series");
    const double s_out = c.scaling_factor_real.empty() ? S.T[1]->noise : c.scaling_factor_real[lout];
    CtPtr res = S.rec(cf, tmax, s_out);
    if (res->level != lout) throw Error(E_STATE, "Chebyshev: level plan and evaluation disagree");
    ct_copy(c, *res, out);
}

// EvalMod of bootstrapping (FIDESlib approxModReductionSparse,
// ApproxModEval.cu:94-130): the Chebyshev series on [-1, 1], the double-angle
// iterations (constants at the tracked scale), then the integer factor `post`.
void eval_mod_reduction(Context& c, const Ciphertext& x, const std::vector<double>& coeffs, int double_angle_its,
                        uint64_t post, Ciphertext& out) {
    eval_chebyshev_series(c, x, coeffs, -1.0, 1.0, out);
    double_angle_at_scale(c, out, double_angle_its);
    eval_mult_int(c, out, post, out);
}

}  // namespace fides

// Chebyshev series sum_k c_k T_k(y), y = (2x - (a + b)) / (b - a), by
// Paterson-Stockmeyer in the Chebyshev basis -- the polynomial step of
// bootstrapping's EvalMod (FIDESlib evalChebyshevSeries,
// src/CKKS/ApproxModEval.cu:370-700, OpenFHE EvalChebyshevSeriesPS), built
// from the engine's ops (with FIXEDMANUAL's explicit rescales under every
// technique but FLEXIBLEAUTOEXT; a FLEXIBLEAUTO result lands on its level's
// ScalingFactorReal):
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
// limbs.  A batch of ciphertexts (the two halves of a bootstrap, or those of
// several bootstraps) is evaluated in lockstep: every product of the series
// is one batched key switch over the whole batch.
// The operation order is this file's own, so results are checked against the
// FP64 value of the series, not residue by residue against the reference.
#include <cmath>

#include "context.hpp"

namespace fides {

namespace {

using CtPtr = std::unique_ptr<Ciphertext>;
using CtVec = std::vector<CtPtr>;
using Scales = std::vector<double>;

std::vector<const Ciphertext*> cptrs(const CtVec& v) {
    std::vector<const Ciphertext*> o;
    for (const auto& p : v) o.push_back(p.get());
    return o;
}

struct Cheb {
    Context& c;
    int B;                     // batch size
    std::vector<CtVec> T;      // T[j][b], j = 1..m
    std::vector<CtVec> G;      // G[t][b] = T_(m 2^t)
    int m = 1;

    CtVec fresh() {
        CtVec v;
        for (int b = 0; b < B; ++b) v.push_back(ct_create(c));
        return v;
    }
    // x[b] * y[b], relinearized and rescaled, one batched key switch
    CtVec mult(const std::vector<const Ciphertext*>& x, const std::vector<const Ciphertext*>& y) {
        CtVec p = fresh();
        if (B == 1) {
            eval_mult(c, *x[0], *y[0], *p[0], true);
        } else {
            std::vector<Ciphertext*> o;
            for (auto& q : p) o.push_back(q.get());
            eval_mult_batch(c, x, y, o, true);
        }
        return p;
    }
    // lower-level copies of a and b (level of the lower one)
    static void align(CtPtr& a, CtPtr& b, Context& c) {
        if (a->level > b->level) drop_to_level(c, *a, b->level);
        if (b->level > a->level) drop_to_level(c, *b, a->level);
    }
    // 2 a b - s (s steered onto the product's scale) or 2 a b - k
    CtVec two_ab_minus(const CtVec& a, const CtVec& b, const CtVec* s, double k) {
        CtVec p = mult(cptrs(a), cptrs(b));
        for (int i = 0; i < B; ++i) {
            eval_add(c, *p[i], *p[i], *p[i]);
            if (s) {
                const Ciphertext& si = *(*s)[i];
                if (si.level <= p[i]->level) throw Error(E_STATE, "Chebyshev: subtrahend has no level to spare");
                CtPtr u = ct_create(c);
                mult_const_to_scale(c, si, p[i]->level + 1, 1.0, p[i]->noise, *u);
                rescale(c, *u);
                eval_sub(c, *p[i], *u, *p[i]);
            } else {
                add_const_at_scale(c, *p[i], -k, *p[i]);
            }
        }
        return p;
    }
    static void trim(std::vector<double>& cf) {
        while (!cf.empty() && cf.back() == 0.0) cf.pop_back();
    }
    // level of base(cf): the lowest level among the T_j it uses, less the rescale
    int base_level(const std::vector<double>& cf) const {
        int lo = T[1][0]->level;
        for (size_t j = 1; j < cf.size(); ++j)
            if (cf[j] != 0.0) lo = std::min(lo, T[j][0]->level);
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
        const int pl = std::min(rec_level(q, t - 1), G[t][0]->level) - 1;
        return std::min(pl, rec_level(r, t - 1));
    }
    // sum_(j<n) cf[j] T_j at scales S (n <= m + 1; T_m taken from the baby steps)
    CtVec base(const std::vector<double>& cf, const Scales& S) {
        const int lo = base_level(cf) + 1;
        CtVec acc = fresh();
        for (int b = 0; b < B; ++b) {
            std::vector<const Ciphertext*> xs;
            std::vector<double> vs;
            for (size_t j = 1; j < cf.size(); ++j)
                if (cf[j] != 0.0) {
                    xs.push_back(T[j][b].get());
                    vs.push_back(cf[j]);
                }
            if (xs.size() > 1)  // one launch for the whole sum
                lincomb_to_scale(c, xs, vs, lo, S[b], *acc[b]);
            else if (xs.size() == 1)
                mult_const_to_scale(c, *xs[0], lo, vs[0], S[b], *acc[b]);
            else  // constant: T_1 * 0 + c_0, as the reference's degree-0 case
                mult_const_to_scale(c, *T[1][b], lo, 0.0, S[b], *acc[b]);
            rescale(c, *acc[b]);
            if (!cf.empty() && cf[0] != 0.0) add_const_at_scale(c, *acc[b], cf[0], *acc[b]);
        }
        return acc;
    }
    // sum_j cf[j] T_j at scales S, for cf.size() <= m 2^(t+1)
    CtVec rec(std::vector<double> cf, int t, const Scales& S) {
        trim(cf);
        if ((int)cf.size() <= m + 1 || t < 0) return base(cf, S);
        const int M = m << t;
        if ((int)cf.size() <= M) return rec(cf, t - 1, S);
        std::vector<double> q, r;
        divide(cf, M, q, r);
        // the quotient's scale makes q T_M land on S after the product's rescale
        const int lvl = std::min(rec_level(q, t - 1), G[t][0]->level);
        Scales sq(B);
        for (int b = 0; b < B; ++b) sq[b] = S[b] * c.mod_reduce_factor[lvl] / G[t][b]->noise;
        CtVec Q = rec(q, t - 1, sq);
        CtVec p = mult(cptrs(Q), cptrs(G[t]));
        Scales sp(B);
        for (int b = 0; b < B; ++b) sp[b] = p[b]->noise;
        CtVec R = rec(r, t - 1, sp);
        for (int b = 0; b < B; ++b) {
            align(p[b], R[b], c);
            eval_add(c, *p[b], *R[b], *p[b]);
        }
        return p;
    }
};

}  // namespace

void eval_chebyshev_series_batch(Context& c, const std::vector<const Ciphertext*>& xs,
                                 const std::vector<double>& coeffs, double a, double b,
                                 const std::vector<Ciphertext*>& outs) {
    const int B = (int)xs.size();
    if (B < 1 || (int)outs.size() != B) throw Error(E_ARG, "batch size mismatch");
    for (const Ciphertext* x : xs) {
        if (!x || x->level < 0) throw Error(E_STATE, "ciphertext not loaded");
        if (x->level != xs[0]->level || x->noise_level != xs[0]->noise_level)
            throw Error(E_STATE, "a Chebyshev batch needs inputs at one level and noise degree");
    }
    const int technique = c.scaling_technique;
    if (technique == FLEXIBLEAUTOEXT) throw Error(E_STATE, "Chebyshev series under FLEXIBLEAUTOEXT is not supported");
    if (coeffs.empty()) throw Error(E_ARG, "no coefficients");
    if (!(b > a)) throw Error(E_ARG, "need a < b");
    int deg = (int)coeffs.size() - 1;
    while (deg > 0 && coeffs[deg] == 0.0) --deg;
    ManualScaling manual(c);
    Cheb S{c, B, {}, {}, 1};
    // y = (2x - (a + b)) / (b - a); no level spent when [a, b] = [-1, 1]
    CtVec y;
    for (int i = 0; i < B; ++i) {
        CtPtr t = ct_create(c);
        ct_copy(c, *xs[i], *t);
        if (t->noise_level == 2) rescale(c, *t);
        if (std::fabs(a + 1.0) > 1e-12 || std::fabs(b - 1.0) > 1e-12) {
            CtPtr u = ct_create(c);
            mult_const_to_scale(c, *t, t->level, 2.0 / (b - a), t->noise, *u);
            rescale(c, *u);
            add_const_at_scale(c, *u, -(a + b) / (b - a), *u);
            t = std::move(u);
        }
        y.push_back(std::move(t));
    }
    // baby steps
    while ((S.m + 1) * (S.m + 1) <= deg + 1 && S.m < 64) S.m <<= 1;
    S.T.resize(S.m + 1);
    S.T[1] = std::move(y);
    for (int i = 2; i <= S.m; ++i) {
        if (i % 2 == 0)
            S.T[i] = S.two_ab_minus(S.T[i / 2], S.T[i / 2], nullptr, 1.0);
        else
            S.T[i] = S.two_ab_minus(S.T[i / 2], S.T[i / 2 + 1], &S.T[1], 0.0);
    }
    // giant steps T_(m 2^t) while m 2^t <= deg
    int tmax = -1;
    {
        CtVec g0;
        for (int i = 0; i < B; ++i) {
            g0.push_back(ct_create(c));
            ct_copy(c, *S.T[S.m][i], *g0.back());
        }
        S.G.push_back(std::move(g0));
    }
    if (S.m <= deg) tmax = 0;
    while ((S.m << (tmax + 1)) <= deg) {
        S.G.push_back(S.two_ab_minus(S.G.back(), S.G.back(), nullptr, 1.0));
        ++tmax;
    }
    std::vector<double> cf(coeffs.begin(), coeffs.begin() + deg + 1);
    const int lout = S.rec_level(cf, tmax);
    if (lout < 0) throw Error(E_LEVEL, "not enough levels for the Chebyshev series");
    const double s_out = c.scaling_factor_real.empty() ? S.T[1][0]->noise : c.scaling_factor_real[lout];
    CtVec res = S.rec(cf, tmax, Scales(B, s_out));
    for (int i = 0; i < B; ++i) {
        if (res[i]->level != lout) throw Error(E_STATE, "Chebyshev: level plan and evaluation disagree");
        // FLEXIBLEAUTO: the result is on its level's factor; the table value
        // replaces the tracked one (equal up to FP64 rounding)
        if (technique == FLEXIBLEAUTO && !c.scaling_factor_real.empty()) res[i]->noise = c.scaling_factor_real[lout];
        ct_copy(c, *res[i], *outs[i]);
    }
}

void eval_chebyshev_series(Context& c, const Ciphertext& x, const std::vector<double>& coeffs, double a, double b,
                           Ciphertext& out) {
    eval_chebyshev_series_batch(c, {&x}, coeffs, a, b, {&out});
}

// EvalMod of bootstrapping (FIDESlib approxModReductionSparse,
// ApproxModEval.cu:94-130, and approxModReduction for the two halves of a
// fully packed bootstrap): the Chebyshev series on [-1, 1], the double-angle
// iterations (constants at the tracked scale), then the integer factor
// `post` -- for a batch in lockstep, every square one batched key switch.
void eval_mod_reduction_batch(Context& c, const std::vector<const Ciphertext*>& xs,
                              const std::vector<double>& coeffs, int double_angle_its, uint64_t post,
                              const std::vector<Ciphertext*>& outs) {
    if (c.scaling_technique == FLEXIBLEAUTOEXT) throw Error(E_STATE, "EvalMod under FLEXIBLEAUTOEXT is not supported");
    ManualScaling manual(c);
    eval_chebyshev_series_batch(c, xs, coeffs, -1.0, 1.0, outs);
    if (double_angle_its < 0) throw Error(E_ARG, "negative iteration count");
    const double two_pi = 2.0 * 3.14159265358979323846;
    const std::vector<const Ciphertext*> in(outs.begin(), outs.end());
    for (int j = 1; j <= double_angle_its; ++j) {
        for (Ciphertext* o : outs)
            if (o->noise_level == 2) rescale(c, *o);
        if (outs.size() == 1)
            eval_square(c, *outs[0], *outs[0], false);
        else
            eval_mult_batch(c, in, in, outs, false);
        for (Ciphertext* o : outs) {
            eval_add(c, *o, *o, *o);
            add_const_at_scale(c, *o, -1.0 / std::pow(two_pi, std::pow(2.0, j - double_angle_its)), *o);
        }
    }
    for (Ciphertext* o : outs) eval_mult_int(c, *o, post, *o);
}

void eval_mod_reduction(Context& c, const Ciphertext& x, const std::vector<double>& coeffs, int double_angle_its,
                        uint64_t post, Ciphertext& out) {
    eval_mod_reduction_batch(c, {&x}, coeffs, double_angle_its, post, {&out});
}

}  // namespace fides

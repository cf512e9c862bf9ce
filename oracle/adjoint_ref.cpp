// FP64 host adjoint of the sphere DEM step -- TEST INFRASTRUCTURE ONLY (loaded by tests/ through pyoracle).
//
// A double-precision forward simulator and its reverse sweep for systems of sphere owners and fixed planes, in the
// world frame, with the contact set frozen per window: at every window start the candidate list is rebuilt by
// brute-force search with the drift margin (sphere pairs with |XA - XB| <= rA + rB + 2 beta, spheres with plane
// distance <= r + 2 beta), and the tangential history of a pair present in both lists carries over, as the
// persistent-contact map does on the device. The per-contact force and its reverse, and the rotation of the angular
// velocity, are the templated statements of the device adjoint (adjoint.h) instantiated with T = double; the
// integration is the extended-Taylor scheme (DEMIntegrationKernels.cu, IntegrationVelPassOnExtendedTaylor.cu):
//     v' = v + h (F/m + g),   X' = X + h (v + h/2 (F/m + g)),   w' = Rot(h (w + h/2 alpha)) (w + h alpha).
//
// Parity pin: the reverse sweep is checked against central finite differences of this file's own forward simulator
// (tests/test_adjoint_ref.py); the forward simulator is checked against the device's float trajectory.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>
#include <vector>

#include "../dem-engine-sbel_amd/csrc/adjoint.h"

namespace {

using deme_adj::V3;
using deme_adj::v3;
using V = V3<double>;

struct Contact {
    uint32_t a, b;  // sphere a; sphere b or plane b
    bool plane;
};

}  // namespace

extern "C" {

// Problem description (all arrays host doubles unless stated).
struct ora_adj_problem {
    uint32_t n;           // spheres (each its own owner)
    const double* X0;     // 3n, interleaved xyz
    const double* V0;     // 3n
    const double* W0;     // 3n world-frame angular velocities
    const double* r;      // n
    const double* m;      // n
    const double* I;      // n
    const uint32_t* mat;  // n
    uint32_t nP;          // planes (fixed)
    const double* planeP;        // 3 nP
    const double* planeN;        // 3 nP, unit normals
    const uint32_t* planeMat;    // nP
    const double* planeMass;     // nP
    uint32_t nMat;
    const double* E;  // nMat each
    const double* nu;
    const double* CoR;
    const double* mu;
    double h, gx, gy, gz, beta;
    int nSteps, window;
};

}  // extern "C"

namespace {

struct Sim {
    const ora_adj_problem& p;
    std::vector<double> E, nu, CoR, mu;  // possibly perturbed copies
    explicit Sim(const ora_adj_problem& q)
        : p(q), E(q.E, q.E + q.nMat), nu(q.nu, q.nu + q.nMat), CoR(q.CoR, q.CoR + q.nMat), mu(q.mu, q.mu + q.nMat) {}

    deme_adj::Mat mat(uint32_t k) const { return deme_adj::Mat{E[k], nu[k], CoR[k], mu[k]}; }

    // brute-force candidate list with the drift margin; uniform grid when large
    std::vector<Contact> detect(const std::vector<double>& X) const {
        std::vector<Contact> out;
        const uint32_t n = p.n;
        double rmax = 0;
        for (uint32_t i = 0; i < n; i++) rmax = std::max(rmax, p.r[i]);
        const double cell = 2 * (rmax + p.beta);
        std::unordered_map<int64_t, std::vector<uint32_t>> grid;
        auto key = [](int64_t i, int64_t j, int64_t k) { return (i * 1048576 + j) * 1048576 + k; };
        auto ci = [&](double x) { return (int64_t)std::floor(x / cell) + 400000; };
        for (uint32_t i = 0; i < n; i++) grid[key(ci(X[3 * i]), ci(X[3 * i + 1]), ci(X[3 * i + 2]))].push_back(i);
        std::vector<std::vector<uint32_t>> partners(n);
        for (uint32_t i = 0; i < n; i++) {
            const int64_t a = ci(X[3 * i]), b = ci(X[3 * i + 1]), c = ci(X[3 * i + 2]);
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        auto it = grid.find(key(a + dx, b + dy, c + dz));
                        if (it == grid.end()) continue;
                        for (uint32_t j : it->second) {
                            if (j <= i) continue;
                            const double d0 = X[3 * i] - X[3 * j], d1 = X[3 * i + 1] - X[3 * j + 1],
                                         d2 = X[3 * i + 2] - X[3 * j + 2];
                            const double lim = p.r[i] + p.r[j] + 2 * p.beta;
                            if (d0 * d0 + d1 * d1 + d2 * d2 <= lim * lim) partners[i].push_back(j);
                        }
                    }
        }
        for (uint32_t i = 0; i < n; i++) {
            for (uint32_t k = 0; k < p.nP; k++) {
                const double d = (X[3 * i] - p.planeP[3 * k]) * p.planeN[3 * k] +
                                 (X[3 * i + 1] - p.planeP[3 * k + 1]) * p.planeN[3 * k + 1] +
                                 (X[3 * i + 2] - p.planeP[3 * k + 2]) * p.planeN[3 * k + 2];
                if (d <= p.r[i] + 2 * p.beta) out.push_back({i, k, true});
            }
            std::sort(partners[i].begin(), partners[i].end());
            for (uint32_t j : partners[i]) out.push_back({i, j, false});
        }
        return out;
    }

    deme_adj::ContactIn<double> contact_in(const Contact& c, const std::vector<double>& X, const std::vector<double>& Vv,
                                           const std::vector<double>& W, V dtIn) const {
        deme_adj::ContactIn<double> in;
        const uint32_t a = c.a, b = c.b;
        for (int k = 0; k < 3; k++) in.XA[k] = X[3 * a + k];
        in.vA = v3(Vv[3 * a], Vv[3 * a + 1], Vv[3 * a + 2]);
        in.wA = v3(W[3 * a], W[3 * a + 1], W[3 * a + 2]);
        in.rA = p.r[a];
        in.mA = p.m[a];
        in.mA_ = mat(p.mat[a]);
        in.dtIn = dtIn;
        in.h = p.h;
        in.plane = c.plane;
        if (!c.plane) {
            for (int k = 0; k < 3; k++) in.XB[k] = X[3 * b + k];
            in.vB = v3(Vv[3 * b], Vv[3 * b + 1], Vv[3 * b + 2]);
            in.wB = v3(W[3 * b], W[3 * b + 1], W[3 * b + 2]);
            in.rB = p.r[b];
            in.mB = p.m[b];
            in.mB_ = mat(p.mat[b]);
            in.nP = v3(0., 0., 0.);
        } else {
            for (int k = 0; k < 3; k++) in.XB[k] = p.planeP[3 * b + k];
            in.vB = in.wB = v3(0., 0., 0.);
            in.rB = 1e15;
            in.mB = p.planeMass[b];
            in.mB_ = mat(p.planeMat[b]);
            in.nP = v3(p.planeN[3 * b], p.planeN[3 * b + 1], p.planeN[3 * b + 2]);
        }
        return in;
    }

    // state at every step start, history in of every step, contact lists per window
    struct Trace {
        std::vector<std::vector<double>> X, Vv, W;          // nSteps + 1
        std::vector<std::vector<Contact>> lists;            // per window
        std::vector<std::vector<int64_t>> map;              // per window: index into the previous window's list or -1
        std::vector<std::vector<V>> hist;                   // per step: history in
    };

    static uint64_t pair_key(const Contact& c) { return ((uint64_t)c.a << 33) | ((uint64_t)c.b << 1) | (c.plane ? 1 : 0); }

    void forward(Trace& T) const {
        const uint32_t n = p.n;
        const double h = p.h;
        T.X.assign(1, std::vector<double>(p.X0, p.X0 + 3 * n));
        T.Vv.assign(1, std::vector<double>(p.V0, p.V0 + 3 * n));
        T.W.assign(1, std::vector<double>(p.W0, p.W0 + 3 * n));
        T.lists.clear();
        T.map.clear();
        T.hist.clear();
        std::vector<V> dt;  // history out of the last step, in the current list's order
        for (int s = 0; s < p.nSteps; s++) {
            if (s % p.window == 0) {
                std::vector<Contact> L = detect(T.X[s]);
                std::vector<int64_t> mp(L.size(), -1);
                std::vector<V> nd(L.size(), v3(0., 0., 0.));
                if (!T.lists.empty()) {
                    std::unordered_map<uint64_t, int64_t> prev;
                    const auto& P = T.lists.back();
                    for (size_t i = 0; i < P.size(); i++) prev[pair_key(P[i])] = (int64_t)i;
                    for (size_t i = 0; i < L.size(); i++) {
                        auto it = prev.find(pair_key(L[i]));
                        if (it != prev.end()) {
                            mp[i] = it->second;
                            nd[i] = dt[it->second];
                        }
                    }
                }
                T.lists.push_back(std::move(L));
                T.map.push_back(std::move(mp));
                dt = std::move(nd);
            }
            const auto& L = T.lists.back();
            T.hist.push_back(dt);
            const auto& X = T.X[s];
            const auto& Vv = T.Vv[s];
            const auto& W = T.W[s];
            std::vector<double> F(3 * n, 0.0), Tq(3 * n, 0.0);
            for (size_t c = 0; c < L.size(); c++) {
                deme_adj::ContactIn<double> in = contact_in(L[c], X, Vv, W, dt[c]);
                deme_adj::ContactTape<double> tp;
                deme_adj::ContactOut<double> o;
                deme_adj::contact_forward(in, tp, o);
                dt[c] = o.dtOut;
                const uint32_t a = L[c].a;
                const V ta = deme_adj::cross(o.armA, o.F);
                F[3 * a] += o.F.x;
                F[3 * a + 1] += o.F.y;
                F[3 * a + 2] += o.F.z;
                Tq[3 * a] += ta.x;
                Tq[3 * a + 1] += ta.y;
                Tq[3 * a + 2] += ta.z;
                if (!L[c].plane) {
                    const uint32_t b = L[c].b;
                    const V tb = deme_adj::cross(o.armB, -o.F);
                    F[3 * b] -= o.F.x;
                    F[3 * b + 1] -= o.F.y;
                    F[3 * b + 2] -= o.F.z;
                    Tq[3 * b] += tb.x;
                    Tq[3 * b + 1] += tb.y;
                    Tq[3 * b + 2] += tb.z;
                }
            }
            std::vector<double> Xn(3 * n), Vn(3 * n), Wn(3 * n);
            const double g[3] = {p.gx, p.gy, p.gz};
            for (uint32_t i = 0; i < n; i++) {
                V al;
                double acc[3];
                for (int k = 0; k < 3; k++) {
                    acc[k] = F[3 * i + k] / p.m[i] + g[k];
                    Xn[3 * i + k] = X[3 * i + k] + h * (Vv[3 * i + k] + 0.5 * h * acc[k]);
                    Vn[3 * i + k] = Vv[3 * i + k] + h * acc[k];
                }
                al = v3(Tq[3 * i] / p.I[i], Tq[3 * i + 1] / p.I[i], Tq[3 * i + 2] / p.I[i]);
                const V w = v3(W[3 * i], W[3 * i + 1], W[3 * i + 2]);
                const V u = w + h * al;
                const V th = h * (w + (0.5 * h) * al);
                const V wn = deme_adj::rot_apply(th, u);
                Wn[3 * i] = wn.x;
                Wn[3 * i + 1] = wn.y;
                Wn[3 * i + 2] = wn.z;
            }
            T.X.push_back(std::move(Xn));
            T.Vv.push_back(std::move(Vn));
            T.W.push_back(std::move(Wn));
        }
    }

    // reverse sweep; gMat: 4 nMat (E, nu, CoR, mu per material)
    void backward(const Trace& T, const double* dLdX, const double* dLdV, double* gV0, double* gMat) const {
        const uint32_t n = p.n;
        const double h = p.h;
        std::vector<double> lX(dLdX ? std::vector<double>(dLdX, dLdX + 3 * n) : std::vector<double>(3 * n, 0.0));
        std::vector<double> lV(dLdV ? std::vector<double>(dLdV, dLdV + 3 * n) : std::vector<double>(3 * n, 0.0));
        std::vector<double> lW(3 * n, 0.0);
        std::vector<double> gm(4 * p.nMat, 0.0);
        int win = (int)T.lists.size() - 1;
        std::vector<V> lDt(T.lists[win].size(), v3(0., 0., 0.));
        for (int s = p.nSteps - 1; s >= 0; s--) {
            const auto& L = T.lists[win];
            const auto& X = T.X[s];
            const auto& Vv = T.Vv[s];
            const auto& W = T.W[s];
            // re-evaluate the step's contacts with tapes, and the accelerations
            std::vector<deme_adj::ContactIn<double>> ins(L.size());
            std::vector<deme_adj::ContactTape<double>> tps(L.size());
            std::vector<deme_adj::ContactOut<double>> outs(L.size());
            std::vector<double> Tq(3 * n, 0.0);
            for (size_t c = 0; c < L.size(); c++) {
                ins[c] = contact_in(L[c], X, Vv, W, T.hist[s][c]);
                deme_adj::contact_forward(ins[c], tps[c], outs[c]);
                const V ta = deme_adj::cross(outs[c].armA, outs[c].F);
                Tq[3 * L[c].a] += ta.x;
                Tq[3 * L[c].a + 1] += ta.y;
                Tq[3 * L[c].a + 2] += ta.z;
                if (!L[c].plane) {
                    const V tb = deme_adj::cross(outs[c].armB, -outs[c].F);
                    Tq[3 * L[c].b] += tb.x;
                    Tq[3 * L[c].b + 1] += tb.y;
                    Tq[3 * L[c].b + 2] += tb.z;
                }
            }
            // integration reverse
            std::vector<double> la(3 * n), lal(3 * n);
            for (uint32_t i = 0; i < n; i++) {
                for (int k = 0; k < 3; k++) {
                    la[3 * i + k] = 0.5 * h * h * lX[3 * i + k] + h * lV[3 * i + k];
                    lV[3 * i + k] += h * lX[3 * i + k];
                }
                const V al = v3(Tq[3 * i] / p.I[i], Tq[3 * i + 1] / p.I[i], Tq[3 * i + 2] / p.I[i]);
                const V w = v3(W[3 * i], W[3 * i + 1], W[3 * i + 2]);
                const V u = w + h * al;
                const V th = h * (w + (0.5 * h) * al);
                V lu, lth;
                deme_adj::rot_reverse(th, u, v3(lW[3 * i], lW[3 * i + 1], lW[3 * i + 2]), lu, lth);
                const V lwn = lu + h * lth;
                const V lalv = h * lu + (0.5 * h * h) * lth;
                lW[3 * i] = lwn.x;
                lW[3 * i + 1] = lwn.y;
                lW[3 * i + 2] = lwn.z;
                lal[3 * i] = lalv.x;
                lal[3 * i + 1] = lalv.y;
                lal[3 * i + 2] = lalv.z;
            }
            // contacts reverse
            std::vector<V> lDtIn(L.size());
            for (size_t c = 0; c < L.size(); c++) {
                const uint32_t a = L[c].a;
                const V ltA = (1.0 / p.I[a]) * v3(lal[3 * a], lal[3 * a + 1], lal[3 * a + 2]);
                V lF = (1.0 / p.m[a]) * v3(la[3 * a], la[3 * a + 1], la[3 * a + 2]);
                deme_adj::acc(lF, deme_adj::cross(ltA, outs[c].armA));
                const V lArmA = deme_adj::cross(outs[c].F, ltA);
                V lArmB = v3(0., 0., 0.);
                if (!L[c].plane) {
                    const uint32_t b = L[c].b;
                    const V ltB = (1.0 / p.I[b]) * v3(lal[3 * b], lal[3 * b + 1], lal[3 * b + 2]);
                    deme_adj::acc(lF, (-1.0 / p.m[b]) * v3(la[3 * b], la[3 * b + 1], la[3 * b + 2]));
                    deme_adj::acc(lF, -deme_adj::cross(ltB, outs[c].armB));
                    lArmB = -deme_adj::cross(outs[c].F, ltB);
                }
                deme_adj::ContactAdj<double> ad;
                deme_adj::contact_reverse(ins[c], tps[c], lF, lArmA, lArmB, lDt[c], ad);
                lDtIn[c] = ad.dtIn;
                const V xs[3] = {ad.XA, ad.vA, ad.wA};
                double* dst[3] = {lX.data(), lV.data(), lW.data()};
                for (int q = 0; q < 3; q++) {
                    dst[q][3 * a] += xs[q].x;
                    dst[q][3 * a + 1] += xs[q].y;
                    dst[q][3 * a + 2] += xs[q].z;
                }
                const uint32_t mA = p.mat[a];
                gm[4 * mA] += ad.gA.E;
                gm[4 * mA + 1] += ad.gA.nu;
                gm[4 * mA + 2] += ad.gA.CoR;
                gm[4 * mA + 3] += ad.gA.mu;
                uint32_t mB;
                if (!L[c].plane) {
                    const uint32_t b = L[c].b;
                    const V ys[3] = {ad.XB, ad.vB, ad.wB};
                    for (int q = 0; q < 3; q++) {
                        dst[q][3 * b] += ys[q].x;
                        dst[q][3 * b + 1] += ys[q].y;
                        dst[q][3 * b + 2] += ys[q].z;
                    }
                    mB = p.mat[b];
                } else {
                    mB = p.planeMat[L[c].b];
                }
                gm[4 * mB] += ad.gB.E;
                gm[4 * mB + 1] += ad.gB.nu;
                gm[4 * mB + 2] += ad.gB.CoR;
                gm[4 * mB + 3] += ad.gB.mu;
            }
            lDt = std::move(lDtIn);
            if (s % p.window == 0 && win > 0) {
                // back through the carry-over of histories onto the previous window's list
                std::vector<V> prev(T.lists[win - 1].size(), v3(0., 0., 0.));
                for (size_t i = 0; i < lDt.size(); i++)
                    if (T.map[win][i] >= 0) prev[T.map[win][i]] = lDt[i];
                lDt = std::move(prev);
                win--;
            }
        }
        if (gV0) std::copy(lV.begin(), lV.end(), gV0);
        if (gMat) std::copy(gm.begin(), gm.end(), gMat);
    }
};

}  // namespace

extern "C" {

// Forward only, with optionally perturbed materials (nullptr: the problem's). Final positions and velocities out.
int ora_adj_forward(const ora_adj_problem* p, const double* E, const double* nu, const double* CoR, const double* mu,
                    double* Xout, double* Vout, uint64_t* nContacts) {
    if (!p || p->nSteps <= 0 || p->window <= 0) return 1;
    Sim s(*p);
    if (E) s.E.assign(E, E + p->nMat);
    if (nu) s.nu.assign(nu, nu + p->nMat);
    if (CoR) s.CoR.assign(CoR, CoR + p->nMat);
    if (mu) s.mu.assign(mu, mu + p->nMat);
    Sim::Trace T;
    s.forward(T);
    if (Xout) std::copy(T.X.back().begin(), T.X.back().end(), Xout);
    if (Vout) std::copy(T.Vv.back().begin(), T.Vv.back().end(), Vout);
    if (nContacts) *nContacts = T.lists.empty() ? 0 : T.lists.back().size();
    return 0;
}

// Forward + reverse: gradients of L with dL/dX_final = dLdX, dL/dV_final = dLdV (3n each, nullable) with respect to the
// initial velocities (3n) and the materials (4 nMat: E, nu, CoR, mu of each).
int ora_adj_gradient(const ora_adj_problem* p, const double* dLdX, const double* dLdV, double* gV0, double* gMat,
                     double* Xout, double* Vout) {
    if (!p || p->nSteps <= 0 || p->window <= 0) return 1;
    Sim s(*p);
    Sim::Trace T;
    s.forward(T);
    if (Xout) std::copy(T.X.back().begin(), T.X.back().end(), Xout);
    if (Vout) std::copy(T.Vv.back().begin(), T.Vv.back().end(), Vout);
    s.backward(T, dLdX, dLdV, gV0, gMat);
    return 0;
}

}  // extern "C"

extern "C" {

// y = Rot(th) u and the adjoints of u and th for a given lY (3 doubles each): the rotation statements checked alone.
void ora_adj_rot(const double* th, const double* u, const double* lY, double* y, double* lU, double* lTh) {
    const V t = v3(th[0], th[1], th[2]), uu = v3(u[0], u[1], u[2]), l = v3(lY[0], lY[1], lY[2]);
    const V r = deme_adj::rot_apply(t, uu);
    V a, b;
    deme_adj::rot_reverse(t, uu, l, a, b);
    y[0] = r.x, y[1] = r.y, y[2] = r.z;
    lU[0] = a.x, lU[1] = a.y, lU[2] = a.z;
    lTh[0] = b.x, lTh[1] = b.y, lTh[2] = b.z;
}

}  // extern "C"

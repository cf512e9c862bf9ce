// Reverse-mode (adjoint) math of one DEM step for sphere owners: the frictional Hertz-Mindlin contact force with
// tangential history (FullHertzianForceModel.cu) and the extended-Taylor integration (DEMIntegrationKernels.cu,
// IntegrationVelPassOnExtendedTaylor.cu), written once for any scalar type T so the device kernels (T = float) and the
// host double-precision reference share the statements.
//
// Model differentiated (the forward simulator's arithmetic as real functions):
//   contact arms from the overlap geometry (checkSpheresOverlap / checkSphereEntityOverlap, plane depth halved);
//   relative velocity at the contact point with world-frame angular velocity (for a sphere R(q)(w_body x R^T a) equals
//   w_world x a); history dt <- project_perp(dt + h v_t); F_n = (k_n d + g_n v_n) n; F_t = -k_t dt - g_t v_t capped at
//   mu |F_n| with dt back-solved; torques a x (+-F); v += h (F/m + g); X += h (v + h/2 (F/m + g)); the body-frame
//   angular update q <- q * exp(h w_pass / 2) in world frame is w' = Rot(h w_pass)(w + h tau / I).
// Positions are treated as continuous (the sub-voxel rounding has zero derivative almost everywhere); the contact set
// is frozen over a window; branches (overlap sign, Coulomb cap, min/max of material pairs) take the active side.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DEME_HD __host__ __device__ __forceinline__
#else
#define DEME_HD inline
#endif

namespace deme_adj {

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
DEME_HD V3<T> v3(T x, T y, T z) {
    return V3<T>{x, y, z};
}
template <typename T>
DEME_HD V3<T> operator+(V3<T> a, V3<T> b) {
    return v3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <typename T>
DEME_HD V3<T> operator-(V3<T> a, V3<T> b) {
    return v3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <typename T>
DEME_HD V3<T> operator-(V3<T> a) {
    return v3(-a.x, -a.y, -a.z);
}
template <typename T>
DEME_HD V3<T> operator*(T s, V3<T> a) {
    return v3(s * a.x, s * a.y, s * a.z);
}
template <typename T>
DEME_HD T dot(V3<T> a, V3<T> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
DEME_HD V3<T> cross(V3<T> a, V3<T> b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T>
DEME_HD T norm(V3<T> a) {
    return sqrt(dot(a, a));
}
template <typename T>
DEME_HD void acc(V3<T>& a, V3<T> b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}

// Material pair (matProxy2ContactParam, DEMHelperKernels.cu:316-339).
struct Mat {
    double E, nu, CoR, mu;
};

template <typename T>
struct MatGrad {
    T E, nu, CoR, mu;
};

constexpr double C1 = 1.825741858350554;  // 2 sqrt(5/6)
constexpr double PI2 = 9.869604401089358;
constexpr double TINY = 1e-12;

// One contact at one step. Inputs: positions (double), velocities and world angular velocities, radii, masses, history
// in, materials. For a plane B: plane point P and unit normal nP, B at rest.
template <typename T>
struct ContactIn {
    double XA[3], XB[3];  // sphere centres; for a plane B, XB = plane point
    V3<T> vA, vB, wA, wB;
    T rA, rB, mA, mB;  // masses of the two bodies as the force model sees them
    V3<T> dtIn;
    Mat mA_, mB_;
    bool plane;
    V3<T> nP;  // plane normal (plane contacts)
    T h;
};

// Everything the reverse sweep needs, recorded by the forward evaluation.
template <typename T>
struct ContactTape {
    bool active;  // overlapping
    T dist, delta, Rs, ms, s, Sn, kn, L, beta, gn, vn, fn, kt, gt, ft, fmax, Es, Gs, CoR, mu, S, Q;
    V3<T> d, n, armA, armB, vrel, vt, dt1, dt2, Ft0, u;
    int branch;  // 0: no tangential force; 1: uncapped; 2: capped
    bool hasMu;
};

template <typename T>
struct ContactOut {
    V3<T> F, armA, armB, dtOut;
};

// Forward evaluation with tape.
template <typename T>
DEME_HD void contact_forward(const ContactIn<T>& in, ContactTape<T>& t, ContactOut<T>& o) {
    const T h = in.h;
    t.active = false;
    o.F = v3<T>(0, 0, 0);
    o.armA = o.armB = v3<T>(0, 0, 0);
    o.dtOut = v3<T>(0, 0, 0);
    if (!in.plane) {
        t.d = v3<T>((T)(in.XA[0] - in.XB[0]), (T)(in.XA[1] - in.XB[1]), (T)(in.XA[2] - in.XB[2]));
        t.dist = norm(t.d);
        t.delta = in.rA + in.rB - t.dist;
        if (!(t.delta >= (T)0) || !(t.dist > (T)0)) return;
        t.n = ((T)1 / t.dist) * t.d;
        t.armA = (-(in.rA - t.delta / (T)2)) * t.n;
        t.armB = (in.rB - t.delta / (T)2) * t.n;
    } else {
        t.n = in.nP;
        t.dist = (T)((in.XA[0] - in.XB[0]) * (double)in.nP.x + (in.XA[1] - in.XB[1]) * (double)in.nP.y +
                     (in.XA[2] - in.XB[2]) * (double)in.nP.z);
        t.delta = (in.rA - t.dist) / (T)2;
        if (!(t.delta >= (T)0)) return;
        t.armA = (-(t.dist + t.delta)) * t.n;
        t.armB = v3<T>(0, 0, 0);
    }
    t.active = true;
    o.armA = t.armA;
    o.armB = t.armB;
    // material pair
    const Mat& A = in.mA_;
    const Mat& B = in.mB_;
    t.S = (T)((1. - A.nu * A.nu) / A.E + (1. - B.nu * B.nu) / B.E);
    t.Es = (T)1 / t.S;
    t.Q = (T)(2. * (2. - A.nu) * (1. + A.nu) / A.E + 2. * (2. - B.nu) * (1. + B.nu) / B.E);
    t.Gs = (T)1 / t.Q;
    t.CoR = (T)(A.CoR < B.CoR ? A.CoR : B.CoR);
    t.mu = (T)(A.mu > B.mu ? A.mu : B.mu);
    // kinematics
    const V3<T> vcA = in.vA + cross(in.wA, t.armA);
    const V3<T> vcB = in.vB + cross(in.wB, t.armB);
    t.vrel = vcA - vcB;
    t.vn = dot(t.vrel, t.n);
    t.vt = t.vrel - t.vn * t.n;
    t.dt1 = in.dtIn + h * t.vt;
    const T p = dot(t.dt1, t.n);
    t.dt2 = t.dt1 - p * t.n;
    // normal force
    t.ms = in.mA * in.mB / (in.mA + in.mB);
    t.Rs = in.rA * in.rB / (in.rA + in.rB);
    t.s = sqrt(t.delta * t.Rs);
    t.Sn = (T)2 * t.Es * t.s;
    t.kn = (T)(2.0 / 3.0) * t.Sn;
    t.L = log(t.CoR < (T)TINY ? (T)TINY : t.CoR);
    t.beta = t.L / sqrt(t.L * t.L + (T)PI2);
    t.gn = (T)C1 * t.beta * sqrt(t.Sn * t.ms);
    t.fn = t.kn * t.delta + t.gn * t.vn;
    o.F = t.fn * t.n;
    // tangential force
    t.hasMu = t.mu > (T)0;
    t.branch = 0;
    o.dtOut = t.dt2;
    if (t.hasMu) {
        t.kt = (T)8 * t.Gs * t.s;
        t.gt = -(T)C1 * t.beta * sqrt(t.ms * t.kt);
        t.Ft0 = (-t.kt) * t.dt2 - t.gt * t.vt;
        t.ft = norm(t.Ft0);
        t.fmax = fabs(t.fn) * t.mu;
        if (t.ft > (T)TINY) {
            if (t.ft > t.fmax) {
                t.branch = 2;
                t.u = ((T)1 / t.ft) * t.Ft0;
                const V3<T> Ft = t.fmax * t.u;
                o.dtOut = ((T)-1 / t.kt) * (Ft + t.gt * t.vt);
                acc(o.F, Ft);
            } else {
                t.branch = 1;
                acc(o.F, t.Ft0);
            }
        }
    }
}

// Reverse sweep: given the adjoints of F, of both arms and of the history out, accumulate the adjoints of the inputs.
template <typename T>
struct ContactAdj {
    V3<T> XA, XB, vA, vB, wA, wB, dtIn;
    MatGrad<T> gA, gB;
};

template <typename T>
DEME_HD void contact_reverse(const ContactIn<T>& in, const ContactTape<T>& t, V3<T> lF, V3<T> lArmA, V3<T> lArmB,
                             V3<T> lDtOut, ContactAdj<T>& a) {
    const T h = in.h;
    a.XA = a.XB = a.vA = a.vB = a.wA = a.wB = a.dtIn = v3<T>(0, 0, 0);
    a.gA = a.gB = MatGrad<T>{0, 0, 0, 0};
    if (!t.active) return;  // no contact: F = 0 and the history is cleared, independent of every input
    V3<T> lN = v3<T>(0, 0, 0), lVt = v3<T>(0, 0, 0), lDt2 = v3<T>(0, 0, 0);
    T lDelta = 0, lFn = 0, lKt = 0, lGt = 0, lMu = 0, lS = 0, lBeta = 0, lEs = 0, lGs = 0;
    // ---- tangential part ----
    if (t.hasMu) {
        V3<T> lFt0 = v3<T>(0, 0, 0);
        if (t.branch == 2) {
            const V3<T> Ft = t.fmax * t.u;
            const V3<T> lFtTot = lF + ((T)-1 / t.kt) * lDtOut;
            // dtOut = (Ft + gt vt) / (-kt)
            const V3<T> dtOut = ((T)-1 / t.kt) * (Ft + t.gt * t.vt);
            lGt += -dot(lDtOut, t.vt) / t.kt;
            acc(lVt, (-t.gt / t.kt) * lDtOut);
            lKt += -dot(lDtOut, dtOut) / t.kt;
            // Ft = fmax u, u = Ft0 / |Ft0|
            const T lFmax = dot(lFtTot, t.u);
            const V3<T> lU = t.fmax * lFtTot;
            lFt0 = ((T)1 / t.ft) * (lU - dot(lU, t.u) * t.u);
            // fmax = |fn| mu
            lFn += (t.fn >= (T)0 ? (T)1 : (T)-1) * t.mu * lFmax;
            lMu += fabs(t.fn) * lFmax;
        } else {
            if (t.branch == 1) lFt0 = lF;
            acc(lDt2, lDtOut);
        }
        if (t.branch != 0) {
            // Ft0 = -kt dt2 - gt vt
            lKt += -dot(lFt0, t.dt2);
            acc(lDt2, (-t.kt) * lFt0);
            lGt += -dot(lFt0, t.vt);
            acc(lVt, (-t.gt) * lFt0);
        }
        // gt = -C1 beta sqrt(ms kt); kt = 8 Gs s
        const T smk = sqrt(t.ms * t.kt);
        lBeta += -(T)C1 * smk * lGt;
        // at zero overlap kt = 0 and dgt/dkt is unbounded; the force and its material derivatives vanish there, and
        // the subgradient 0 is taken
        if (smk > (T)0) lKt += -(T)C1 * t.beta * t.ms / ((T)2 * smk) * lGt;
        lGs += (T)8 * t.s * lKt;
        lS += (T)8 * t.Gs * lKt;
    } else {
        acc(lDt2, lDtOut);
    }
    // ---- normal part: F_n = fn n ----
    lFn += dot(lF, t.n);
    acc(lN, t.fn * lF);
    // fn = kn delta + gn vn
    const T lKn = t.delta * lFn;
    lDelta += t.kn * lFn;
    const T lGn = t.vn * lFn;
    T lVn = t.gn * lFn;
    // gn = C1 beta sqrt(Sn ms), kn = 2/3 Sn, Sn = 2 Es s
    const T sqSm = sqrt(t.Sn * t.ms);
    lBeta += (T)C1 * sqSm * lGn;
    T lSn = (T)(2.0 / 3.0) * lKn;
    if (sqSm > (T)0) lSn += (T)C1 * t.beta * t.ms / ((T)2 * sqSm) * lGn;  // zero overlap: as for gt above
    lEs += (T)2 * t.s * lSn;
    lS += (T)2 * t.Es * lSn;
    // beta = L / sqrt(L^2 + pi^2), L = log(CoR)
    const T den = t.L * t.L + (T)PI2;
    const T lL = (T)PI2 / (den * sqrt(den)) * lBeta;
    const T lCoR = t.CoR < (T)TINY ? (T)0 : lL / t.CoR;
    // s = sqrt(delta Rs)
    if (t.s > (T)0) lDelta += t.Rs / ((T)2 * t.s) * lS;
    // ---- history: dt2 = dt1 - (dt1.n) n, dt1 = dtIn + h vt ----
    const T p = dot(t.dt1, t.n);
    V3<T> lDt1 = lDt2;
    const T lP = -dot(lDt2, t.n);
    acc(lN, (-p) * lDt2);
    acc(lDt1, lP * t.n);
    acc(lN, lP * t.dt1);
    a.dtIn = lDt1;
    acc(lVt, h * lDt1);
    // ---- relative velocity: vt = vrel - vn n, vn = vrel.n ----
    V3<T> lVrel = lVt;
    lVn += -dot(lVt, t.n);
    acc(lN, (-t.vn) * lVt);
    acc(lVrel, lVn * t.n);
    acc(lN, lVn * t.vrel);
    // vcA = vA + wA x armA, vcB = vB + wB x armB, vrel = vcA - vcB
    a.vA = lVrel;
    a.wA = cross(t.armA, lVrel);
    acc(lArmA, cross(lVrel, in.wA));
    if (!in.plane) {
        a.vB = -lVrel;
        a.wB = cross(t.armB, -lVrel);
        acc(lArmB, cross(-lVrel, in.wB));
    }
    // ---- geometry ----
    if (!in.plane) {
        // armA = -(rA - delta/2) n, armB = (rB - delta/2) n
        acc(lN, (-(in.rA - t.delta / (T)2)) * lArmA);
        lDelta += dot(lArmA, t.n) / (T)2;
        acc(lN, (in.rB - t.delta / (T)2) * lArmB);
        lDelta += -dot(lArmB, t.n) / (T)2;
        // delta = rA + rB - dist, n = d / dist
        const T lDist = -lDelta;
        V3<T> lD = ((T)1 / t.dist) * (lN - dot(lN, t.n) * t.n);
        acc(lD, lDist * t.n);
        a.XA = lD;
        a.XB = -lD;
    } else {
        // armA = -(dist + delta) nP, delta = (rA - dist) / 2, dist = (XA - P).nP
        const T lAn = dot(lArmA, t.n);
        T lDist = -lAn;
        lDelta += -lAn;
        lDist += -lDelta / (T)2;
        a.XA = lDist * t.n;
    }
    // ---- materials: Es = 1/S, Gs = 1/Q, CoR = min, mu = max ----
    const Mat& A = in.mA_;
    const Mat& B = in.mB_;
    const T lSs = -lEs / (t.S * t.S);
    const T lQ = -lGs / (t.Q * t.Q);
    a.gA.E += lSs * (T)(-(1. - A.nu * A.nu) / (A.E * A.E)) + lQ * (T)(-2. * (2. - A.nu) * (1. + A.nu) / (A.E * A.E));
    a.gA.nu += lSs * (T)(-2. * A.nu / A.E) + lQ * (T)(2. * (1. - 2. * A.nu) / A.E);
    a.gB.E += lSs * (T)(-(1. - B.nu * B.nu) / (B.E * B.E)) + lQ * (T)(-2. * (2. - B.nu) * (1. + B.nu) / (B.E * B.E));
    a.gB.nu += lSs * (T)(-2. * B.nu / B.E) + lQ * (T)(2. * (1. - 2. * B.nu) / B.E);
    if (A.CoR <= B.CoR) a.gA.CoR += lCoR;
    else a.gB.CoR += lCoR;
    if (A.mu >= B.mu) a.gA.mu += lMu;
    else a.gB.mu += lMu;
}

// ---- rotation of a vector by the rotation vector th: Rot(th) u (Rodrigues) and its adjoints ----
template <typename T>
DEME_HD void rot_coeffs(T phi, T& s, T& c, T& ds, T& dc) {
    // s = sin(phi)/phi, c = (1 - cos(phi))/phi^2, ds = s'(phi)/phi, dc = c'(phi)/phi
    if (phi < (T)1e-2) {
        const T p2 = phi * phi;
        s = (T)1 - p2 / (T)6 + p2 * p2 / (T)120;
        c = (T)0.5 - p2 / (T)24 + p2 * p2 / (T)720;
        ds = -(T)1 / (T)3 + p2 / (T)30;
        dc = -(T)1 / (T)12 + p2 / (T)180;
    } else {
        const T sn = sin(phi), cs = cos(phi);
        s = sn / phi;
        c = ((T)1 - cs) / (phi * phi);
        ds = (cs * phi - sn) / (phi * phi * phi);
        dc = (sn * phi - (T)2 * ((T)1 - cs)) / (phi * phi * phi * phi);
    }
}

template <typename T>
DEME_HD V3<T> rot_apply(V3<T> th, V3<T> u) {
    T s, c, ds, dc;
    rot_coeffs(norm(th), s, c, ds, dc);
    const V3<T> w = cross(th, u);
    return u + s * w + c * cross(th, w);
}

// adjoints of u and th for y = Rot(th) u, given lY
template <typename T>
DEME_HD void rot_reverse(V3<T> th, V3<T> u, V3<T> lY, V3<T>& lU, V3<T>& lTh) {
    const T phi = norm(th);
    T s, c, ds, dc;
    rot_coeffs(phi, s, c, ds, dc);
    const V3<T> w = cross(th, u);   // th x u
    const V3<T> w2 = cross(th, w);  // th x (th x u)
    const V3<T> lxt = cross(lY, th);
    lU = lY + s * lxt + c * cross(lxt, th);
    lTh = s * cross(u, lY) + c * cross(w, lY) + c * cross(u, cross(lY, th)) + (ds * dot(w, lY)) * th +
          (dc * dot(w2, lY)) * th;
}

}  // namespace deme_adj

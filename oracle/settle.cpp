// Oracle: closed-form rigid-water constraints (SETTLE). Test infrastructure
// only (see oracle.hpp).
//
// The reference plans SETTLE (Miyamoto & Kollman, J. Comput. Chem. 13, 952
// (1992)) as a replacement of RigidWater's SHAKE/RATTLE that must reproduce
// it to machine precision (constrain/rigid_water.hpp:20-24); it is absent
// from the reference's sources. This is a restatement of the published
// method from its defining conditions, derived here:
//
// Positions. SHAKE's fixed point (rigid_water.cpp:61-116) moves site i by
// D_i = (1/m_i) sum_k g_k b_k, b_k the reference (start-of-step) bonds. Such
// displacements (a) lie in the plane of the reference molecule, (b) keep the
// centre of mass, (c) carry no angular momentum about it: sum m_i r0_i x D_i
// = sum_k g_k b_k x b_k = 0. Conversely, in-plane displacements meeting (b)
// and (c) span a 3-dimensional space, that of the three g_k. So the solution
// is the rigid canonical triangle placed at the unconstrained centre of mass
// and rotated by the R that meets (a) and (c):
//  * frame of the reference molecule: ez = normal(b0 x c0), ex = unit(H1->H2),
//    ey = ez x ex; canonical sites (relative to the centre of mass) O (0, ra),
//    H1 (-rc, -rb), H2 (rc, -rb), rc = dHH/2, ra + rb = h = sqrt(dOH^2 - rc^2),
//    mO ra = 2 mH rb;
//  * (a) fixes the third row of R from the z components u_iz of the
//    unconstrained sites: R_zy = u_Oz / ra, R_zx = (u_2z - u_1z) / (2 rc),
//    R_zz = +sqrt(1 - R_zx^2 - R_zy^2) (the branch that is the identity for
//    a zero step);
//  * R = Rz(theta) M, with M the rotation about ez x r3 taking ez to the third
//    row r3 = (s1, s2, s3) (Rodrigues):
//      M = [[1 - s1^2 k, -s1 s2 k, -s1], [-s1 s2 k, 1 - s2^2 k, -s2], [s1, s2, s3]],
//      k = 1 / (1 + s3);
//  * (c) reads Sxx R_yx - Syy R_xy = T, Sxx = 2 mH rc^2, Syy = mO ra^2 +
//    2 mH rb^2 (the cross moment is zero), T = sum m_i (p_ix u_iy - p_iy u_ix);
//    with alpha = Sxx M_xx + Syy M_yy and beta = (Sxx - Syy) M_xy it is
//    alpha sin(theta) + beta cos(theta) = T, whose root near zero is
//      sin = (T alpha - Q beta) / rho2, cos = (Q alpha + T beta) / rho2,
//      rho2 = alpha^2 + beta^2, Q = sqrt(rho2 - T^2).
// Each site moves by D_i = R p_i - u_i (frame components, back to the lab
// frame), added to the stored position as SHAKE adds its corrections, so
// molecules that straddle the box keep their images.
//
// Velocities. RATTLE (rigid_water.cpp:118-161) is Gauss-Seidel on the
// linear system for the three bond multipliers mu: with current bonds r_l
// (minimum image), v_l the relative velocity of bond l, and
//   A_ll = (1/m_i + 1/m_j) r_l.r_l,  A_(01)(02) = r01.r02 / mO,
//   A_(01)(12) = -r01.r12 / mH,      A_(02)(12) = r02.r12 / mH,
// A mu = (r_l.v_l); then v_i -= mu_l r_l / m_i, v_j += mu_l r_l / m_j for
// every bond l = (i, j). The 3 x 3 system is solved by Cramer's rule.
#include <string>

#include "oracle.hpp"

namespace oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;

V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 scale(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 unit(V3 a) { return scale(a, 1.0 / std::sqrt(dot3(a, a))); }

void check_molecules(const State& s) {
    if (s.n() % 3 != 0) throw std::runtime_error("Settle: atom count is not a multiple of 3");
    for (std::size_t m0 = 0; m0 < s.n(); m0 += 3)
        if (s.m[m0 + 1] != s.m[m0 + 2])
            throw std::runtime_error("Settle: the two hydrogen masses of a molecule differ");
}
}  // namespace

Settle::Settle(double oh_dist, double hoh_deg) : oh(oh_dist), hh(0.0) {
    if (!(oh_dist > 0.0)) throw std::invalid_argument("Settle: oh distance must be positive");
    if (!(hoh_deg > 0.0) || !(hoh_deg < 180.0))
        throw std::invalid_argument("Settle: angle must lie in (0, 180)");
    hh = 2.0 * oh_dist * std::sin(0.5 * hoh_deg * kPi / 180.0);
}

int Settle::positions(State& s, const std::vector<V3>& reference) const {
    check_molecules(s);
    if (reference.size() != s.n()) throw std::runtime_error("Settle: reference size mismatch");
    const double rc = 0.5 * hh;
    const double h = std::sqrt(oh * oh - rc * rc);
    for (std::size_t a = 0; a < s.n(); a += 3) {
        const double mo = s.m[a], mh = s.m[a + 1];
        const double mt = mo + 2.0 * mh;
        const double ra = 2.0 * mh * h / mt, rb = h - ra;
        // reference frame
        const V3 b0 = min_image(s.L, sub(reference[a + 1], reference[a]));
        const V3 c0 = min_image(s.L, sub(reference[a + 2], reference[a]));
        const V3 nz = cross(b0, c0);
        const double nn = dot3(nz, nz);
        if (!(nn > 0.0)) throw std::runtime_error("Settle: reference molecule is degenerate");
        const V3 ez = scale(nz, 1.0 / std::sqrt(nn));
        const V3 ex = unit(sub(c0, b0));
        const V3 ey = cross(ez, ex);
        // unconstrained sites relative to their centre of mass
        const V3 b1 = min_image(s.L, sub(s.r(a + 1), s.r(a)));
        const V3 c1 = min_image(s.L, sub(s.r(a + 2), s.r(a)));
        const V3 d = scale(add(scale(b1, mh), scale(c1, mh)), 1.0 / mt);
        const V3 u[3] = {scale(d, -1.0), sub(b1, d), sub(c1, d)};
        double ux[3], uy[3], uz[3];
        for (int k = 0; k < 3; ++k) {
            ux[k] = dot3(u[k], ex);
            uy[k] = dot3(u[k], ey);
            uz[k] = dot3(u[k], ez);
        }
        const double s2 = uz[0] / ra, s1 = (uz[2] - uz[1]) / (2.0 * rc);
        const double q = 1.0 - s1 * s1 - s2 * s2;
        if (!(q > 0.0))
            throw std::runtime_error("Settle: molecule rotated out of its plane by more than 90 degrees");
        const double s3 = std::sqrt(q);
        const double k = 1.0 / (1.0 + s3);
        const double mxx = 1.0 - s1 * s1 * k, myy = 1.0 - s2 * s2 * k, mxy = -s1 * s2 * k;
        const double sxx = 2.0 * mh * rc * rc, syy = mo * ra * ra + 2.0 * mh * rb * rb;
        const double T = -mo * ra * ux[0] + mh * rc * (uy[2] - uy[1]) + mh * rb * (ux[1] + ux[2]);
        const double al = sxx * mxx + syy * myy, be = (sxx - syy) * mxy;
        const double rho2 = al * al + be * be;
        const double disc = rho2 - T * T;
        if (!(disc >= 0.0)) throw std::runtime_error("Settle: in-plane rotation has no solution");
        const double Q = std::sqrt(disc);
        const double st = (T * al - Q * be) / rho2, ct = (Q * al + T * be) / rho2;
        // rows of R = Rz(theta) M (columns x, y; z is not needed: p_z = 0)
        const double r1x = ct * mxx - st * mxy, r1y = ct * mxy - st * myy;
        const double r2x = st * mxx + ct * mxy, r2y = st * mxy + ct * myy;
        const double px[3] = {0.0, -rc, rc}, py[3] = {ra, -rb, -rb};
        for (int i = 0; i < 3; ++i) {
            const double nx = r1x * px[i] + r1y * py[i];
            const double ny = r2x * px[i] + r2y * py[i];
            const double nzc = s1 * px[i] + s2 * py[i];
            const double dx = nx - ux[i], dy = ny - uy[i], dz = nzc - uz[i];
            const V3 D = add(add(scale(ex, dx), scale(ey, dy)), scale(ez, dz));
            s.set_r(a + i, add(s.r(a + i), D));
        }
    }
    return 1;
}

int Settle::velocities(State& s) const {
    check_molecules(s);
    for (std::size_t a = 0; a < s.n(); a += 3) {
        const double mo = s.m[a], mh = s.m[a + 1];
        const V3 r01 = min_image(s.L, sub(s.r(a), s.r(a + 1)));
        const V3 r02 = min_image(s.L, sub(s.r(a), s.r(a + 2)));
        const V3 r12 = min_image(s.L, sub(s.r(a + 1), s.r(a + 2)));
        const V3 v0 = s.v(a), v1 = s.v(a + 1), v2 = s.v(a + 2);
        const double ioh = 1.0 / mo + 1.0 / mh;
        const double a00 = ioh * dot3(r01, r01), a11 = ioh * dot3(r02, r02),
                     a22 = (2.0 / mh) * dot3(r12, r12);
        const double a01 = dot3(r01, r02) / mo, a02 = -dot3(r01, r12) / mh, a12 = dot3(r02, r12) / mh;
        const double b0 = dot3(r01, sub(v0, v1)), b1 = dot3(r02, sub(v0, v2)), b2 = dot3(r12, sub(v1, v2));
        // cofactors of the symmetric matrix
        const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
        const double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
        const double det = a00 * c00 + a01 * c01 + a02 * c02;
        if (!(std::abs(det) > 0.0)) throw std::runtime_error("Settle: velocity constraints are singular");
        const double inv = 1.0 / det;
        const double mu0 = (c00 * b0 + c01 * b1 + c02 * b2) * inv;
        const double mu1 = (c01 * b0 + c11 * b1 + c12 * b2) * inv;
        const double mu2 = (c02 * b0 + c12 * b1 + c22 * b2) * inv;
        const V3 g01 = scale(r01, mu0), g02 = scale(r02, mu1), g12 = scale(r12, mu2);
        s.set_v(a, sub(v0, scale(add(g01, g02), 1.0 / mo)));
        s.set_v(a + 1, add(v1, scale(sub(g01, g12), 1.0 / mh)));
        s.set_v(a + 2, add(v2, scale(add(g02, g12), 1.0 / mh)));
    }
    return 1;
}

}  // namespace oracle

// CPU oracle for the FWI acoustic solver path -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference's Fortran, templated on float and
// double.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only as the checker.  The product (libfwi_hip.so) never
// links it and never falls back to it.
//
// The restatement deliberately keeps the reference's structure (full-array
// derivative passes, one array set per CPML region, per-region local
// derivatives, paste and rotate), which is a different formulation from the
// fused single-pass GPU kernel, so agreement between the two checks both.
//
// Arithmetic order follows the Fortran statements term by term (left-to-right
// evaluation, unary minus of a product applied to the whole product, integer
// constants converted to the working precision) and the file is compiled with
// -ffp-contract=off so no multiply-add is fused.
//
// Pinning (what checks each function, besides the GPU comparison):
//   fd_coef           : reference finite_difference_coef_ (oracle/_ref) + Fornberg
//                       closed form and polynomial exactness (tests/test_oracle_units.py)
//   space derivatives : reference space_deriv2x_/2y_/x_/y_ bitwise + polynomial exactness
//   pml_coeffs        : reference pml_coeffs_ bitwise + closed-form profile checks
//   solver            : reference solver_acoustic_ss_ (free surface on and off, FP32 and
//                       FP64 builds) bitwise / to roundoff, and committed golden vectors
//   forward_simple    : reference solver_forward_simple_ (built at -O0)
//   misfit_l2 core    : reference misfit_l2_ bitwise
//   trace normalisation (Misfit.f90:28-61): brute force on small traces; the
//                       reference Misfit_function cannot be built unmodified here
//                       (its module needs an `iargc` declaration), so this part is
//                       pinned by closed-form tests only.
//   gradient_post     : closed form (shot sum, scale, water mask) in the tests.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <class T> struct Lit;
template <> struct Lit<float> {
    static float pi() { return 3.14159265f; }   // solver_PML_coeffs.f90:16 (REAL(4) literal)
    static float refl() { return 0.00001f; }    // solver_PML_coeffs.f90:17
};
template <> struct Lit<double> {
    static double pi() { return 3.14159265; }   // same literals under -fdefault-real-8
    static double refl() { return 0.00001; }
};

// Column-major 2-D array (rows fastest), 0-based indices, like Fortran A(ny,nx).
template <class T> struct Arr {
    int ny = 0, nx = 0;
    std::vector<T> v;
    Arr() = default;
    Arr(int ny_, int nx_) : ny(ny_), nx(nx_), v((size_t)ny_ * nx_, T(0)) {}
    T& operator()(int j, int i) { return v[(size_t)i * ny + j]; }
    const T& operator()(int j, int i) const { return v[(size_t)i * ny + j]; }
    void zero() { std::fill(v.begin(), v.end(), T(0)); }
};

// solver_FD_coeffs.f90:1-73.  num=1: first derivative, num=2: second
// derivative (centre weight halved).  Reproduces the table as written,
// including acc=1 coef(2)=1/2 (integer division, 0) and acc=5 num=1 zeros.
template <class T> void fd_coef(int num, int acc, T* coef) {
    if (num == 1) {
        if (acc == 1) { coef[0] = T(0); coef[1] = T(0); }
        if (acc == 3) { coef[0] = T(0); coef[1] = T(3) / T(4); coef[2] = T(-3) / T(20); coef[3] = T(1) / T(60); }
        if (acc == 4) { coef[0] = T(0); coef[1] = T(4) / T(5); coef[2] = T(-1) / T(5); coef[3] = T(4) / T(105); coef[4] = T(-1) / T(280); }
        if (acc == 5) { for (int i = 0; i < 6; ++i) coef[i] = T(0); }
    }
    if (num == 2) {
        if (acc == 3) { coef[0] = T(-49) / T(18); coef[1] = T(3) / T(2); coef[2] = T(-3) / T(20); coef[3] = T(1) / T(90); }
        if (acc == 4) { coef[0] = T(-205) / T(72); coef[1] = T(8) / T(5); coef[2] = T(-1) / T(5); coef[3] = T(8) / T(315); coef[4] = T(-1) / T(560); }
        if (acc == 5) { coef[0] = T(-5269) / T(1800); coef[1] = T(5) / T(3); coef[2] = T(-5) / T(21); coef[3] = T(5) / T(126); coef[4] = T(-5) / T(1008); coef[5] = T(1) / T(3150); }
        coef[0] = coef[0] / T(2);
    }
}

// solver_space_derivs.f90.  kind: 0 = d2/dx2 (:1), 1 = d2/dy2 (:28),
// 2 = d/dx (:60), 3 = d/dy (:99).  x is the column index, y the row index.
// Valid positions acc..n-1-acc (0-based); the outer acc lines copy the
// nearest computed line.
template <class T> void space_deriv(int kind, int acc, const T* coef, const Arr<T>& p, Arr<T>& out) {
    const int nyt = p.ny, nxt = p.nx;
    if (kind == 0 || kind == 2) {
        const int j1 = acc, j2 = nxt - acc - 1;
        for (int i = j1; i <= j2; ++i)
            for (int j = 0; j < nyt; ++j) {
                T s = T(0);
                for (int m = 0; m <= acc; ++m) {
                    if (kind == 0) s = s + coef[m] * (p(j, i + m) + p(j, i - m));
                    else           s = s + (p(j, i + m) - p(j, i - m)) * coef[m];
                }
                out(j, i) = s;
            }
        for (int i = 0; i < j1; ++i)
            for (int j = 0; j < nyt; ++j) out(j, i) = out(j, j1);
        for (int i = j2 + 1; i < nxt; ++i)
            for (int j = 0; j < nyt; ++j) out(j, i) = out(j, j2);
    } else {
        const int j1 = acc, j2 = nyt - acc - 1;
        for (int i = 0; i < nxt; ++i)
            for (int j = j1; j <= j2; ++j) {
                T s = T(0);
                for (int m = 0; m <= acc; ++m) {
                    if (kind == 1) s = s + (p(j + m, i) + p(j - m, i)) * coef[m];
                    else           s = s + (p(j + m, i) - p(j - m, i)) * coef[m];
                }
                out(j, i) = s;
            }
        for (int i = 0; i < nxt; ++i) {
            for (int j = 0; j < j1; ++j) out(j, i) = out(j1, i);
            for (int j = j2 + 1; j < nyt; ++j) out(j, i) = out(j2, i);
        }
    }
}

// solver_PML_coeffs.f90:1-75.  fA/fB have dn entries (index 0 = interior
// side); the corner matrices are dn x dn column-major.
template <class T> void pml_coeffs(int acc, int dPML, T dt, T d, int dn, T* fA, T* fB,
                                   T* ldA, T* rdA, T* luA, T* ruA, T* ldB, T* rdB, T* luB, T* ruB) {
    const T pi = Lit<T>::pi();
    const T refl = Lit<T>::refl();
    const T wPML = T(dPML - 1) * d;
    std::vector<T> yPML(dn, T(0)), alpha(dn, T(0)), sigma(dn, T(0));
    for (int i = 0; i < dn; ++i) { fA[i] = T(0); fB[i] = T(0); }
    for (int i1 = acc + 1; i1 <= dPML + acc; ++i1) {       // Fortran index i1
        const int i = i1 - 1;
        yPML[i] = T(i1 - (acc + 1)) * d;
        alpha[i] = (T(30) * pi) * (T(1) - yPML[i] / wPML);
        const T q = yPML[i] / wPML;
        sigma[i] = (-((T(6000) * std::log(refl)) / (T(2) * wPML))) * (q * q);
    }
    for (int i1 = acc + 1; i1 <= dPML + acc; ++i1) {
        const int i = i1 - 1;
        fA[i] = std::exp(-((sigma[i] + alpha[i]) * dt));
        fB[i] = ((fA[i] - T(1)) * sigma[i]) / (sigma[i] + alpha[i]);
    }
    for (int i1 = dPML + acc + 1; i1 <= dn; ++i1) {
        fA[i1 - 1] = fA[dPML + acc - 1];
        fB[i1 - 1] = fB[dPML + acc - 1];
    }
    auto at = [dn](T* M, int j1, int i1) -> T& { return M[(size_t)(i1 - 1) * dn + (j1 - 1)]; };
    for (int i1 = 1; i1 <= dn; ++i1)
        for (int j1 = 1; j1 <= i1; ++j1) { at(rdA, j1, i1) = fA[i1 - 1]; at(rdB, j1, i1) = fB[i1 - 1]; }
    for (int i1 = 1; i1 <= dn; ++i1)
        for (int j1 = i1 + 1; j1 <= dn; ++j1) { at(rdA, j1, i1) = at(rdA, i1, j1); at(rdB, j1, i1) = at(rdB, i1, j1); }
    for (int i1 = 1; i1 <= dn; ++i1)
        for (int j1 = 1; j1 <= dn - i1 + 1; ++j1) { at(ldA, j1, i1) = fA[dn - i1]; at(ldB, j1, i1) = fB[dn - i1]; }
    for (int ix = 1; ix <= dn - 1; ++ix) {
        const int j1 = 2 + (ix - 1);
        for (int i1 = dn - (ix - 1); i1 <= dn; ++i1) { at(ldA, j1, i1) = fA[j1 - 1]; at(ldB, j1, i1) = fB[j1 - 1]; }
    }
    for (int j1 = 1; j1 <= dn; ++j1)
        for (int i1 = 1; i1 <= dn; ++i1) {
            at(luA, j1, i1) = at(ldA, dn - j1 + 1, i1);
            at(luB, j1, i1) = at(ldB, dn - j1 + 1, i1);
            at(ruA, j1, i1) = at(rdA, dn - j1 + 1, i1);
            at(ruB, j1, i1) = at(rdB, dn - j1 + 1, i1);
        }
}

// One CPML region: origin (y0,x0) in the padded grid, size nvy x nvx, and the
// memory variables of solver_acoustic_SS.f90:55-85 (phi/chi and their _1 copies)
// for up to two directions.
template <class T> struct Region {
    int y0, x0, nvy, nvx;
    Arr<T> phix, phix1, chix, chix1, phiy, phiy1, chiy, chiy1, p3;
    Region(int y0_, int x0_, int nvy_, int nvx_)
        : y0(y0_), x0(x0_), nvy(nvy_), nvx(nvx_),
          phix(nvy_, nvx_), phix1(nvy_, nvx_), chix(nvy_, nvx_), chix1(nvy_, nvx_),
          phiy(nvy_, nvx_), phiy1(nvy_, nvx_), chiy(nvy_, nvx_), chiy1(nvy_, nvx_), p3(nvy_, nvx_) {}
};

// pml_matrices_x / _y / _xycorner (solver_PML_matrices.f90:1,42,83).
// dir: 0 = x, 1 = y.  fac(j,i) returns the (A,B) factor index pair source.
// The local derivative is taken over the region's own phi array (clamped at
// the region edges) before phi is overwritten; phi equals phi_1 on entry.
template <class T, class FacA, class FacB>
void pml_region(int dir, int acc, const T* coef1, Region<T>& R, const Arr<T>& deriv, const Arr<T>& deriv2,
                FacA facA, FacB facB, Arr<T>& prop) {
    Arr<T>& phi = dir == 0 ? R.phix : R.phiy;
    Arr<T>& phi1 = dir == 0 ? R.phix1 : R.phiy1;
    Arr<T>& chi = dir == 0 ? R.chix : R.chiy;
    Arr<T>& chi1 = dir == 0 ? R.chix1 : R.chiy1;
    Arr<T> dphi(R.nvy, R.nvx);
    space_deriv(dir == 0 ? 2 : 3, acc, coef1, phi, dphi);
    for (int i = 0; i < R.nvx; ++i)
        for (int j = 0; j < R.nvy; ++j) {
            const T A = facA(j, i), B = facB(j, i);
            phi(j, i) = A * phi1(j, i) + B * deriv(R.y0 + j, R.x0 + i);
            chi(j, i) = A * chi1(j, i) + B * (deriv2(R.y0 + j, R.x0 + i) + dphi(j, i));
        }
    for (int i = 0; i < R.nvx; ++i)
        for (int j = 0; j < R.nvy; ++j) prop(j, i) = dphi(j, i) + chi(j, i);
}

struct Opts {
    int dpml;          // CPML thickness in cells (20 in solver_acoustic_SS, 50 in solver_forward_simple)
    int free_surface;  // 1: top ghost rows held at zero (reference default); 0: CPML on top too
};

// solver_acoustic_SS.f90:9-645.  Arguments follow the reference routine:
// 1-based model positions, Fortran column-major arrays S(nt,NumSou),
// Sadj(nt,NumRec), ctmp(ny,nx), kernel(ny,nx), Fields(ny,nx*(nt_Store+1)),
// Rec(nt,NumRec).  iprop=1 forward with snapshots, 2 adjoint with the imaging
// condition, anything else forward only.  Returns 0, or a negative code for
// inputs on which the reference would index out of bounds.
template <class T>
int solver(const Opts& o, int iprop, int NumSou, int NumRec, const int* nxSou, const int* nySou,
           const int* nxRec, const int* nyRec, const T* S, const T* Sadj, T d, T dt, int ny, int nx,
           const T* ctmp, int store_snap, int nt_Store, int nt, T* kernel, T* Fields, T* Rec) {
    const int acc = 3;
    const int dn = o.dpml + 2 * acc;
    const int nyt = ny + 2 * dn, nxt = nx + 2 * dn;
    if (ny < 1 || nx < 1 || nt < 2) return -1;
    std::fill(Rec, Rec + (size_t)nt * NumRec, T(0));           // :92

    // injection cells (:104-118), 0-based padded coordinates
    std::vector<int> sx, sy, rx, ry;
    if (iprop != 2) {
        for (int j = 0; j < NumSou; ++j) { sx.push_back(dn + nxSou[j] - 1); sy.push_back(dn + nySou[j] - 1); }
        for (int r = 0; r < NumRec; ++r) { rx.push_back(dn + nxRec[r] - 1); ry.push_back(dn + nyRec[r] - 1); }
    } else {
        for (int r = 0; r < NumRec; ++r) { sx.push_back(dn + nxRec[r] - 1); sy.push_back(dn + nyRec[r] - 1); }
    }
    for (size_t j = 0; j < sx.size(); ++j)
        if (sx[j] < 0 || sx[j] >= nxt || sy[j] < 0 || sy[j] >= nyt) return -2;
    for (size_t j = 0; j < rx.size(); ++j)
        if (rx[j] < 0 || rx[j] >= nxt || ry[j] < 0 || ry[j] >= nyt) return -2;

    // velocity padded by nearest-edge extension (:271-290), then c2 (:292)
    Arr<T> c2(nyt, nxt);
    for (int i = 0; i < nxt; ++i)
        for (int j = 0; j < nyt; ++j) {
            const int jm = std::min(std::max(j - dn, 0), ny - 1);
            const int im = std::min(std::max(i - dn, 0), nx - 1);
            const T c = ctmp[(size_t)im * ny + jm];
            c2(j, i) = ((c * c) * (dt * dt)) / (d * d);
        }

    std::vector<T> fA(dn), fB(dn);
    std::vector<T> ldA(dn * dn), rdA(dn * dn), luA(dn * dn), ruA(dn * dn), ldB(dn * dn), rdB(dn * dn), luB(dn * dn), ruB(dn * dn);
    pml_coeffs<T>(acc, o.dpml, dt, d, dn, fA.data(), fB.data(), ldA.data(), rdA.data(), luA.data(), ruA.data(),
                  ldB.data(), rdB.data(), luB.data(), ruB.data());
    T coef1[4] = {0, 0, 0, 0}, coef2[4] = {0, 0, 0, 0};
    fd_coef<T>(1, acc, coef1);
    fd_coef<T>(2, acc, coef2);

    Arr<T> src(nyt, nxt), p1(nyt, nxt), p(nyt, nxt), p3(nyt, nxt);
    Arr<T> d2x(nyt, nxt), d2y(nyt, nxt), d1x(nyt, nxt), d1y(nyt, nxt);

    // regions (:160-209): strips and corners, padded origins
    Region<T> xr(dn, dn + nx, ny, dn), xl(dn, 0, ny, dn);
    Region<T> yd(dn + ny, dn, dn, nx), yu(0, dn, dn, nx);
    Region<T> lu(0, 0, dn, dn), ru(0, dn + nx, dn, dn), ld(dn + ny, 0, dn, dn), rd(dn + ny, dn + nx, dn, dn);
    Arr<T> xprop(ny, dn), yprop(dn, nx), cprop(dn, dn);

    auto vecA = [](const std::vector<T>& f) { return f.data(); };
    const T* FA = vecA(fA);
    const T* FB = vecA(fB);
    auto mat = [dn](const std::vector<T>& M) { return [&M, dn](int j, int i) { return M[(size_t)i * dn + j]; }; };

    // strip p update (:349-350 etc.): p3 = -p1 + 2 p + c2 (d2x + d2y + prop)
    auto strip_update = [&](Region<T>& R, const Arr<T>& prop) {
        for (int i = 0; i < R.nvx; ++i)
            for (int j = 0; j < R.nvy; ++j) {
                const int Y = R.y0 + j, X = R.x0 + i;
                R.p3(j, i) = (-p1(Y, X) + T(2) * p(Y, X)) + c2(Y, X) * ((d2x(Y, X) + d2y(Y, X)) + prop(j, i));
            }
    };
    // corner y part added in a second statement (:380,389,402,411)
    auto corner_add = [&](Region<T>& R, const Arr<T>& prop) {
        for (int i = 0; i < R.nvx; ++i)
            for (int j = 0; j < R.nvy; ++j) {
                const int Y = R.y0 + j, X = R.x0 + i;
                R.p3(j, i) = R.p3(j, i) + c2(Y, X) * prop(j, i);
            }
    };
    auto rotate = [](Region<T>& R) { R.phix1 = R.phix; R.chix1 = R.chix; R.phiy1 = R.phiy; R.chiy1 = R.chiy; };
    auto paste = [&](const Region<T>& R) {
        for (int i = 0; i < R.nvx; ++i)
            for (int j = 0; j < R.nvy; ++j) p3(R.y0 + j, R.x0 + i) = R.p3(j, i);
    };

    int l = 1;
    int snap = (iprop == 1) ? 0 : (iprop == 2 ? nt_Store : 0);

    for (int k = 2; k <= nt; ++k) {
        l = l + 1;
        if (iprop != 2) {
            for (int j = 0; j < NumSou; ++j) src(sy[j], sx[j]) = S[(size_t)j * nt + (k - 2)];
        } else {
            for (int j = 0; j < NumRec; ++j) src(sy[j], sx[j]) = Sadj[(size_t)j * nt + (k - 2)];
        }
        space_deriv(0, acc, coef2, p, d2x);
        space_deriv(1, acc, coef2, p, d2y);
        space_deriv(2, acc, coef1, p, d1x);
        space_deriv(3, acc, coef1, p, d1y);

        // x strips (:347-355): xr uses factor(i), xl uses factor(dn-i+1)
        pml_region(0, acc, coef1, xr, d1x, d2x, [&](int, int i) { return FA[i]; }, [&](int, int i) { return FB[i]; }, xprop);
        strip_update(xr, xprop);
        pml_region(0, acc, coef1, xl, d1x, d2x, [&](int, int i) { return FA[dn - 1 - i]; }, [&](int, int i) { return FB[dn - 1 - i]; }, xprop);
        strip_update(xl, xprop);
        // bottom strip (:359-362)
        pml_region(1, acc, coef1, yd, d1y, d2y, [&](int j, int) { return FA[j]; }, [&](int j, int) { return FB[j]; }, yprop);
        strip_update(yd, yprop);
        if (o.free_surface == 0) {
            // top strip and top corners (:366-389)
            pml_region(1, acc, coef1, yu, d1y, d2y, [&](int j, int) { return FA[dn - 1 - j]; }, [&](int j, int) { return FB[dn - 1 - j]; }, yprop);
            strip_update(yu, yprop);
            pml_region(0, acc, coef1, lu, d1x, d2x, mat(luA), mat(luB), cprop);
            strip_update(lu, cprop);
            pml_region(1, acc, coef1, lu, d1y, d2y, mat(luA), mat(luB), cprop);
            corner_add(lu, cprop);
            pml_region(0, acc, coef1, ru, d1x, d2x, mat(ruA), mat(ruB), cprop);
            strip_update(ru, cprop);
            pml_region(1, acc, coef1, ru, d1y, d2y, mat(ruA), mat(ruB), cprop);
            corner_add(ru, cprop);
        }
        // bottom corners (:395-411)
        pml_region(0, acc, coef1, ld, d1x, d2x, mat(ldA), mat(ldB), cprop);
        strip_update(ld, cprop);
        pml_region(1, acc, coef1, ld, d1y, d2y, mat(ldA), mat(ldB), cprop);
        corner_add(ld, cprop);
        pml_region(0, acc, coef1, rd, d1x, d2x, mat(rdA), mat(rdB), cprop);
        strip_update(rd, cprop);
        pml_region(1, acc, coef1, rd, d1y, d2y, mat(rdA), mat(rdB), cprop);
        corner_add(rd, cprop);

        if (o.free_surface == 1) {
            // :416-429 -- zero the first model row, then mirror rows of p3 (which
            // still hold the previous step's pasted values) into the top regions
            for (int i = 0; i < nxt; ++i) p3(dn, i) = T(0);
            for (int j = 0; j < dn; ++j) {
                const int Ysrc = dn - j;   // Fortran dny2+1+(1-j1) with j1=j+1
                for (int i = 0; i < nx; ++i) yu.p3(j, i) = p3(Ysrc, dn + i);
                for (int i = 0; i < dn; ++i) { lu.p3(j, i) = p3(Ysrc, i); ru.p3(j, i) = p3(Ysrc, dn + nx + i); }
            }
        }

        // interior (:433-434)
        for (int i = dn; i < dn + nx; ++i)
            for (int j = dn; j < dn + ny; ++j)
                p3(j, i) = (-p1(j, i) + T(2) * p(j, i)) + ((d2x(j, i) + d2y(j, i)) + src(j, i)) * c2(j, i);

        if (iprop == 1 && l == store_snap) {                        // :438-454
            snap = snap + 1;
            if (snap > nt_Store) return -3;
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    Fields[((size_t)snap * nx + i) * ny + j] = ((p3(dn + j, dn + i) - p1(dn + j, dn + i)) * T(0.5)) / dt;
            l = 0;
        }
        if (iprop == 2 && l == store_snap) {                        // :457-473
            snap = snap - 1;
            if (snap < 0 || snap > nt_Store) return -3;
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j) {
                    const T f = Fields[((size_t)snap * nx + i) * ny + j];
                    const T k2 = -(((T(0.5) * (p3(dn + j, dn + i) - p1(dn + j, dn + i))) * f) * T(store_snap));
                    kernel[(size_t)i * ny + j] = kernel[(size_t)i * ny + j] + k2;
                }
            l = 0;
        }

        // paste regions into p3 (:477-486)
        paste(yu); paste(lu); paste(ru); paste(yd); paste(ld); paste(rd); paste(xl); paste(xr);
        // rotate (:491-547)
        p1 = p;
        p = p3;
        rotate(yu); rotate(yd); rotate(xl); rotate(xr); rotate(lu); rotate(ru); rotate(ld); rotate(rd);
        if (iprop != 2)
            for (int r = 0; r < NumRec; ++r) Rec[(size_t)r * nt + (k - 2)] = p3(ry[r], rx[r]);   // :551-555
    }
    for (int r = 0; r < NumRec; ++r) Rec[(size_t)r * nt + (nt - 1)] = Rec[(size_t)r * nt + (nt - 2)];   // :588
    return 0;
}

// Misfit.f90:28-61 (per-trace max-abs normalisation) then Misfit_L2.f90:8-14.
// Misfit is accumulated sequentially in the working precision, as the
// reference does.  misfit is set (the reference zeroes it at Misfit.f90:67).
template <class T>
void misfit_l2(const T* synth, const T* obs, int nrec, int nt, T dt, T* misfit, T* adj) {
    std::vector<T> DS((size_t)nrec * nt, T(0)), DT((size_t)nrec * nt, T(0));
    for (int j = 0; j < nrec; ++j) {
        const T* s = synth + (size_t)j * nt;
        const T* o = obs + (size_t)j * nt;
        bool nz = false;
        for (int k = 0; k < nt; ++k) if (o[k] != T(0)) { nz = true; break; }
        if (!nz) continue;
        T maxS = T(0), maxT = T(0);
        for (int k = 0; k < nt; ++k) { maxS = std::max(maxS, std::fabs(s[k])); maxT = std::max(maxT, std::fabs(o[k])); }
        if (maxS != T(0) && maxT != T(0))
            for (int k = 0; k < nt; ++k) { DS[(size_t)j * nt + k] = s[k] / maxS; DT[(size_t)j * nt + k] = o[k] / maxT; }
    }
    T m = T(0);
    for (int j = 0; j < nrec; ++j)
        for (int k = 0; k < nt; ++k) {
            const T df = DS[(size_t)j * nt + k] - DT[(size_t)j * nt + k];
            m = m + (T(0.5) * dt) * (df * df);
            const size_t kr = (size_t)j * nt + (nt - 1 - k);
            adj[(size_t)j * nt + k] = -(DS[kr] - DT[kr]);
        }
    *misfit = m;
}

}  // namespace

extern "C" {

struct oracle_opts { int dpml; int free_surface; };

#define ORACLE_SOLVER(T, SUF)                                                                               \
    int oracle_solver_##SUF(const oracle_opts* o, int iprop, int NumSou, int NumRec, const int* nxSou,     \
                            const int* nySou, const int* nxRec, const int* nyRec, const T* S, const T* Sadj,\
                            T d, T dt, int ny, int nx, const T* ctmp, int store_snap, int nt_Store, int nt, \
                            T* kernel, T* Fields, T* Rec) {                                                 \
        Opts op{o->dpml, o->free_surface};                                                                  \
        return solver<T>(op, iprop, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, Sadj, d, dt, ny, nx,     \
                         ctmp, store_snap, nt_Store, nt, kernel, Fields, Rec);                              \
    }                                                                                                       \
    void oracle_fd_coef_##SUF(int num, int acc, T* coef) { fd_coef<T>(num, acc, coef); }                    \
    void oracle_space_deriv_##SUF(int kind, int acc, const T* coef, int nyt, int nxt, const T* p, T* out) { \
        Arr<T> a(nyt, nxt), b(nyt, nxt);                                                                    \
        std::memcpy(a.v.data(), p, sizeof(T) * (size_t)nyt * nxt);                                          \
        space_deriv<T>(kind, acc, coef, a, b);                                                              \
        std::memcpy(out, b.v.data(), sizeof(T) * (size_t)nyt * nxt);                                        \
    }                                                                                                       \
    void oracle_pml_coeffs_##SUF(int acc, int dPML, T dt, T d, int dn, T* fA, T* fB, T* ldA, T* rdA,       \
                                 T* luA, T* ruA, T* ldB, T* rdB, T* luB, T* ruB) {                          \
        pml_coeffs<T>(acc, dPML, dt, d, dn, fA, fB, ldA, rdA, luA, ruA, ldB, rdB, luB, ruB);                \
    }                                                                                                       \
    void oracle_misfit_l2_##SUF(const T* synth, const T* obs, int nrec, int nt, T dt, T* misfit, T* adj) {  \
        misfit_l2<T>(synth, obs, nrec, nt, dt, misfit, adj);                                                \
    }

ORACLE_SOLVER(float, f32)
ORACLE_SOLVER(double, f64)

// solver_forward_simple.f90:4-542: the same scheme with dPML=50, one source
// injected at one cell, free surface on, no snapshots.
int oracle_forward_simple_f32(int NumRec, int nxSou, int nySou, const int* nxRec, const int* nyRec, const float* S,
                              float d, float dt, int ny, int nx, const float* ctmp, int nt, float* Rec) {
    Opts op{50, 1};
    return solver<float>(op, 3, 1, NumRec, &nxSou, &nySou, nxRec, nyRec, S, nullptr, d, dt, ny, nx, ctmp, 1, 0, nt,
                         nullptr, nullptr, Rec);
}

// Shot sum of per-shot imaging kernels into the model gradient, in double:
// grad(y, x0_s + x) += scale * kernel_s(y, x) for every shot s, then the water
// rows y < water_rows[x] are zeroed (calculate_misfit_gradient.f90:339-351,
// with the shot sum done in FP64 instead of an FP32 MPI_REDUCE).
// Arrays are column-major: grad(ny, nx_model), kernel_s(ny, w_s).
void oracle_gradient_post(int ny, int nx_model, int nshots, const int* x0, const int* w, const float* const* kernels,
                          double scale, const int* water_rows, double* grad) {
    for (int s = 0; s < nshots; ++s)
        for (int i = 0; i < w[s]; ++i)
            for (int j = 0; j < ny; ++j)
                grad[(size_t)(x0[s] + i) * ny + j] += scale * (double)kernels[s][(size_t)i * ny + j];
    if (water_rows)
        for (int i = 0; i < nx_model; ++i)
            for (int j = 0; j < std::min(water_rows[i], ny); ++j) grad[(size_t)i * ny + j] = 0.0;
}

}  // extern "C"

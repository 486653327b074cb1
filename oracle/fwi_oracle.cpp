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

// ---------------------------------------------------------------------------
// Exact discrete adjoint (no counterpart in the reference, whose adjoint
// re-runs the forward propagator on time-reversed residuals).
//
// ExactGrid holds the padded grid of one shot, cell-wise: c2 and the CPML
// factor table indexed by the 1-based distance m = max(dx, dy) into the
// frame (m = 0 inside), the same per-cell formulation as the GPU kernel, so
// exact_forward<float> reproduces solver<float> bit for bit.
//
// The adjoint is the transpose of one leapfrog step
//   p_k = 2 p_{k-1} - p_{k-2} + c2 (D2x p_{k-1} + D2y p_{k-1} + X + Y + s_k)
//   X = Dbx phi_{k-1} + chi_k,  phi_k = A phi_{k-1} + B D1x p_{k-1},
//   chi_k = A chi_{k-1} + B (D2x p_{k-1} + Dbx phi_{k-1})          (x band; y alike)
// where D2x, D1x are the clamped derivatives (outer 3 columns copy column 3
// / NX-4) and Dbx the band-local first derivative (clamped to band columns
// 3..dn-4).  A clamped operator is S C (stencil S on the computed lines, C
// copying the clamp line), so its transpose is C^T S^T: fold the outer lines
// onto the clamp line, then apply the mirrored stencil.  Carried backwards:
//   a_k  (adjoint of p_k),  Xt_k = c2 a_k + A Xt_{k+1} (of chi_k),
//   Pt_k = A Pt_{k+1} + Dbx^T u_{k+1} (of phi_k),  u = g + B Xt, g = c2 a,
//   a_k = e_k + 2 a_{k+1} - a_{k+2} + D2x^T t1x + D1x^T t3x + (y alike),
//   t1x = g + [band] B Xt,  t3x = [band] B Pt    (all at level k+1).
// The gradient with respect to c2 is sum_k a_k (p_k - 2 p_{k-1} + p_{k-2})/c2,
// folded onto the model cells the padded c2 copies (nearest-edge extension)
// and chained through c2 = vp^2 dt^2 / d^2.

template <class T> struct ExactGrid {
    int ny, nx, dn, NY, NX, fs;
    std::vector<T> c2;            // [y][x]
    std::vector<T> fA, fB;        // [0..dn] by 1-based distance, 0 inside
    ExactGrid(const Opts& o, int ny_, int nx_, const T* vp, T d, T dt)
        : ny(ny_), nx(nx_), dn(o.dpml + 6), NY(ny_ + 2 * (o.dpml + 6)), NX(nx_ + 2 * (o.dpml + 6)),
          fs(o.free_surface), c2((size_t)NY * NX), fA(dn + 1, T(0)), fB(dn + 1, T(0)) {
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) {
                const int jm = std::min(std::max(y - dn, 0), ny - 1), im = std::min(std::max(x - dn, 0), nx - 1);
                const T c = vp[(size_t)im * ny + jm];
                c2[(size_t)y * NX + x] = ((c * c) * (dt * dt)) / (d * d);
            }
        std::vector<T> a(dn), b(dn), m8((size_t)dn * dn * 8);
        T* M = m8.data();
        const size_t q = (size_t)dn * dn;
        pml_coeffs<T>(3, o.dpml, dt, d, dn, a.data(), b.data(), M, M + q, M + 2 * q, M + 3 * q, M + 4 * q,
                      M + 5 * q, M + 6 * q, M + 7 * q);
        for (int m = 1; m <= dn; ++m) { fA[m] = a[m - 1]; fB[m] = b[m - 1]; }
    }
    int dxv(int x) const { return x < dn ? dn - x : (x >= NX - dn ? x - (NX - dn) + 1 : 0); }
    int dyv(int y) const {
        if (y < dn) return fs ? 0 : dn - y;
        return y >= NY - dn ? y - (NY - dn) + 1 : 0;
    }
    bool upd(int y) const { return !fs || y >= dn; }
    size_t at(int y, int x) const { return (size_t)y * NX + x; }
};

template <class T> struct Stencil {
    T c1[4], c2[4];
    Stencil() {
        c1[0] = T(0); c1[1] = T(3) / T(4); c1[2] = T(-3) / T(20); c1[3] = T(1) / T(60);
        c2[0] = (T(-49) / T(18)) / T(2); c2[1] = T(3) / T(2); c2[2] = T(-3) / T(20); c2[3] = T(1) / T(90);
    }
    // second difference from values v[-3..3] (v points at the centre), as dd2
    T d2(const T* v, ptrdiff_t s) const {
        T r = c2[0] * (v[0] + v[0]);
        r = r + c2[1] * (v[s] + v[-s]);
        r = r + c2[2] * (v[2 * s] + v[-2 * s]);
        r = r + c2[3] * (v[3 * s] + v[-3 * s]);
        return r;
    }
    T d1(const T* v, ptrdiff_t s) const {
        T r = (v[s] - v[-s]) * c1[1];
        r = r + (v[2 * s] - v[-2 * s]) * c1[2];
        r = r + (v[3 * s] - v[-3 * s]) * c1[3];
        return r;
    }
    // weights of the operators as linear maps: out(j) = sum_m w[m+3] in(j+m)
    T w2(int m) const { return m == 0 ? c2[0] * T(2) : c2[std::abs(m)]; }
    T w1(int m) const { return m == 0 ? T(0) : (m > 0 ? c1[m] : -c1[-m]); }
};

// Injection and receiver cells in padded coordinates (model cells only for
// injection, last listed source wins on a shared cell).
struct ExactAcq {
    std::vector<int> src_cell, src_of_cell_tr;   // per source: padded cell or -1; trace injected there
    std::vector<int> rec_cell;
};

template <class T>
ExactAcq exact_acq(const ExactGrid<T>& g, int ns, const int* nxs, const int* nys, int nr, const int* nxr,
                   const int* nyr) {
    ExactAcq a;
    a.src_cell.assign(ns, -1);
    a.src_of_cell_tr.assign(ns, -1);
    std::vector<int> last((size_t)g.NY * g.NX, -1);
    for (int j = 0; j < ns; ++j)
        if (nxs[j] >= 1 && nxs[j] <= g.nx && nys[j] >= 1 && nys[j] <= g.ny) {
            const int c = (int)g.at(g.dn + nys[j] - 1, g.dn + nxs[j] - 1);
            a.src_cell[j] = c;
            last[c] = j;
        }
    for (int j = 0; j < ns; ++j)
        if (a.src_cell[j] >= 0) a.src_of_cell_tr[j] = last[a.src_cell[j]] == j ? j : -1;
    for (int r = 0; r < nr; ++r) a.rec_cell.push_back((int)g.at(g.dn + nyr[r] - 1, g.dn + nxr[r] - 1));
    return a;
}

// Forward modelling, cell-wise.  rec(nt, nr) column-major as Rec_point;
// hist (optional) receives p at every time level 0..nt, [k][y][x].
template <class T>
void exact_forward(const ExactGrid<T>& g, const ExactAcq& acq, int nt, const T* S, T* rec, T* hist) {
    const Stencil<T> st;
    const int NY = g.NY, NX = g.NX, dn = g.dn;
    const size_t n = (size_t)NY * NX;
    const int ns = (int)acq.src_cell.size(), nr = (int)acq.rec_cell.size();
    std::vector<T> p0(n, T(0)), p1(n, T(0)), pn(n, T(0));
    std::vector<T> phx[2] = {std::vector<T>((size_t)2 * NY * dn, T(0)), std::vector<T>((size_t)2 * NY * dn, T(0))};
    std::vector<T> chx((size_t)2 * NY * dn, T(0));
    std::vector<T> phy[2] = {std::vector<T>((size_t)2 * dn * NX, T(0)), std::vector<T>((size_t)2 * dn * NX, T(0))};
    std::vector<T> chy((size_t)2 * dn * NX, T(0));
    std::vector<T> src(n, T(0));
    std::vector<char> inj(n, 0);
    for (int j = 0; j < ns; ++j)
        if (acq.src_of_cell_tr[j] >= 0) inj[acq.src_cell[j]] = 1;
    std::fill(rec, rec + (size_t)nt * nr, T(0));
    if (hist) std::fill(hist, hist + 2 * n, T(0));
    for (int k = 2; k <= nt; ++k) {
        const int o = k & 1;
        for (int j = 0; j < ns; ++j)
            if (acq.src_of_cell_tr[j] >= 0) src[acq.src_cell[j]] = S[(size_t)j * nt + (k - 2)];
        for (int y = 0; y < NY; ++y) {
            if (!g.upd(y)) continue;
            const int yc = std::min(std::max(y, 3), NY - 4);
            const int dyv = g.dyv(y), yband = y < dn ? 0 : 1, rb = y < dn ? y : y - (NY - dn);
            for (int x = 0; x < NX; ++x) {
                const size_t i = g.at(y, x);
                const int xc = std::min(std::max(x, 3), NX - 4);
                const T d2x = st.d2(&p1[g.at(y, xc)], 1);
                const T d2y = st.d2(&p1[g.at(yc, x)], NX);
                const T base = (-p0[i]) + T(2) * p1[i];
                const T lap = d2x + d2y;
                const int dxv = g.dxv(x);
                const T c2 = g.c2[i];
                T pp;
                if (dxv == 0 && dyv == 0) {
                    T qq = lap;
                    if (inj[i]) qq = qq + src[i];
                    pp = base + qq * c2;
                } else {
                    const int m = std::max(dxv, dyv);
                    const T A = g.fA[m], B = g.fB[m];
                    T propx = T(0), propy = T(0);
                    if (dxv > 0) {
                        const int band = x < dn ? 0 : 1, b = x < dn ? x : x - (NX - dn);
                        const int bc = std::min(std::max(b, 3), dn - 4);
                        const size_t row = ((size_t)band * NY + y) * dn;
                        const T d1x = st.d1(&p1[g.at(y, xc)], 1);
                        const T dph = st.d1(&phx[o][row + bc], 1);
                        const T phn = A * phx[o][row + b] + B * d1x;
                        const T chn = A * chx[row + b] + B * (d2x + dph);
                        phx[o ^ 1][row + b] = phn;
                        chx[row + b] = chn;
                        propx = dph + chn;
                    }
                    if (dyv > 0) {
                        const int rc = std::min(std::max(rb, 3), dn - 4);
                        const size_t col = (size_t)yband * dn * NX + x;
                        const T d1y = st.d1(&p1[g.at(yc, x)], NX);
                        const T dph = st.d1(&phy[o][col + (size_t)rc * NX], NX);
                        const T phn = A * phy[o][col + (size_t)rb * NX] + B * d1y;
                        const T chn = A * chy[col + (size_t)rb * NX] + B * (d2y + dph);
                        phy[o ^ 1][col + (size_t)rb * NX] = phn;
                        chy[col + (size_t)rb * NX] = chn;
                        propy = dph + chn;
                    }
                    if (dxv > 0 && dyv > 0) pp = (base + c2 * (lap + propx)) + c2 * propy;
                    else if (dxv > 0) pp = base + c2 * (lap + propx);
                    else pp = base + c2 * (lap + propy);
                }
                pn[i] = pp;
            }
        }
        for (int r = 0; r < nr; ++r) rec[(size_t)r * nt + (k - 2)] = pn[acq.rec_cell[r]];
        if (hist) std::copy(pn.begin(), pn.end(), hist + (size_t)k * n);
        std::swap(p0, p1);
        std::swap(p1, pn);
    }
    for (int r = 0; r < nr; ++r) rec[(size_t)r * nt + (nt - 1)] = rec[(size_t)r * nt + (nt - 2)];
}

// Exact adjoint.  e(nt, nr): the derivative of the objective with respect to
// each recorded sample.  srcadj(nt, ns) (optional) receives the derivative
// with respect to each source sample; grad_vp(ny, nx) (optional, needs hist)
// receives the derivative with respect to the model velocity.
template <class T>
void exact_adjoint(const ExactGrid<T>& g, const ExactAcq& acq, int nt, const T* e, const T* hist, T d, T dt,
                   const T* vp, T* srcadj, double* grad_vp) {
    const Stencil<T> st;
    const int NY = g.NY, NX = g.NX, dn = g.dn;
    const size_t n = (size_t)NY * NX;
    const int ns = (int)acq.src_cell.size(), nr = (int)acq.rec_cell.size();
    std::vector<T> a2(n, T(0)), a1(n, T(0)), an(n, T(0)), rte(n, T(0));
    std::vector<T> t1x(n), t1y(n), t3x(n), t3y(n);
    std::vector<T> Xx[2], Px[2], Xy[2], Py[2];
    for (int b = 0; b < 2; ++b) {
        Xx[b].assign((size_t)2 * NY * dn, T(0)); Px[b].assign((size_t)2 * NY * dn, T(0));
        Xy[b].assign((size_t)2 * dn * NX, T(0)); Py[b].assign((size_t)2 * dn * NX, T(0));
    }
    std::vector<double> gc2(n, 0.0);
    if (srcadj) std::fill(srcadj, srcadj + (size_t)nt * ns, T(0));
    auto xband = [&](int x, int& band, int& b) { band = x < dn ? 0 : 1; b = x < dn ? x : x - (NX - dn); };
    auto yband = [&](int y, int& band, int& rb) { band = y < dn ? 0 : 1; rb = y < dn ? y : y - (NY - dn); };
    int cur = 0;   // Xx[cur] etc. hold level k+1
    for (int k = nt; k >= 2; --k) {
        // level k+1 intermediates (zero at k = nt)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) {
                const size_t i = g.at(y, x);
                const T gk = g.c2[i] * a1[i];
                t1x[i] = gk; t1y[i] = gk; t3x[i] = T(0); t3y[i] = T(0);
                if (!g.upd(y)) continue;
                const int dxv = g.dxv(x), dyv = g.dyv(y), m = std::max(dxv, dyv);
                const T B = g.fB[m];
                if (dxv > 0) {
                    int band, b; xband(x, band, b);
                    const size_t q = ((size_t)band * NY + y) * dn + b;
                    t1x[i] = gk + B * Xx[cur][q];
                    t3x[i] = B * Px[cur][q];
                }
                if (dyv > 0) {
                    int band, rb; yband(y, band, rb);
                    const size_t q = (size_t)band * dn * NX + (size_t)rb * NX + x;
                    t1y[i] = gk + B * Xy[cur][q];
                    t3y[i] = B * Py[cur][q];
                }
            }
        // receiver term
        std::fill(rte.begin(), rte.end(), T(0));
        for (int r = 0; r < nr; ++r) {
            T v = e[(size_t)r * nt + (k - 2)];
            if (k == nt) v = v + e[(size_t)r * nt + (nt - 1)];
            rte[acq.rec_cell[r]] += v;
        }
        // folded line value of a clamped operator's input at computed line j
        auto foldx = [&](const std::vector<T>& t, int y, int j) -> T {
            if (j < 3 || j > NX - 4) return T(0);
            T v = t[g.at(y, j)];
            if (j == 3) v = v + t[g.at(y, 0)] + t[g.at(y, 1)] + t[g.at(y, 2)];
            if (j == NX - 4) v = v + t[g.at(y, NX - 3)] + t[g.at(y, NX - 2)] + t[g.at(y, NX - 1)];
            return v;
        };
        auto foldy = [&](const std::vector<T>& t, int j, int x) -> T {
            if (j < 3 || j > NY - 4) return T(0);
            T v = t[g.at(j, x)];
            if (j == 3) v = v + t[g.at(0, x)] + t[g.at(1, x)] + t[g.at(2, x)];
            if (j == NY - 4) v = v + t[g.at(NY - 3, x)] + t[g.at(NY - 2, x)] + t[g.at(NY - 1, x)];
            return v;
        };
        for (int y = 0; y < NY; ++y) {
            if (!g.upd(y)) continue;
            for (int x = 0; x < NX; ++x) {
                const size_t i = g.at(y, x);
                T v = rte[i] + T(2) * a1[i] - a2[i];
                for (int m = -3; m <= 3; ++m) {
                    v = v + st.w2(m) * foldx(t1x, y, x - m) + st.w1(m) * foldx(t3x, y, x - m);
                    v = v + st.w2(m) * foldy(t1y, y - m, x) + st.w1(m) * foldy(t3y, y - m, x);
                }
                an[i] = v;
            }
        }
        // memory-variable adjoints at level k
        const int nx_ = cur ^ 1;
        for (int y = 0; y < NY; ++y) {
            if (!g.upd(y)) continue;
            for (int x = 0; x < NX; ++x) {
                const int dxv = g.dxv(x), dyv = g.dyv(y), m = std::max(dxv, dyv);
                const T A = g.fA[m];
                const size_t i = g.at(y, x);
                if (dxv > 0) {
                    int band, b; xband(x, band, b);
                    const size_t row = ((size_t)band * NY + y) * dn;
                    // Dbx^T u: u = t1x on band cells of this row, folded onto band columns 3..dn-4
                    auto U = [&](int j) -> T {
                        if (j < 3 || j > dn - 4) return T(0);
                        auto ub = [&](int bb) { return t1x[g.at(y, band == 0 ? bb : bb + (NX - dn))]; };
                        T v = ub(j);
                        if (j == 3) v = v + ub(0) + ub(1) + ub(2);
                        if (j == dn - 4) v = v + ub(dn - 3) + ub(dn - 2) + ub(dn - 1);
                        return v;
                    };
                    T dtu = T(0);
                    for (int mm = -3; mm <= 3; ++mm) dtu = dtu + st.w1(mm) * U(b - mm);
                    Px[nx_][row + b] = A * Px[cur][row + b] + dtu;
                    Xx[nx_][row + b] = g.c2[i] * an[i] + A * Xx[cur][row + b];
                }
                if (dyv > 0) {
                    int band, rb; yband(y, band, rb);
                    const size_t col = (size_t)band * dn * NX + x;
                    auto U = [&](int j) -> T {
                        if (j < 3 || j > dn - 4) return T(0);
                        auto ub = [&](int bb) { return t1y[g.at(band == 0 ? bb : bb + (NY - dn), x)]; };
                        T v = ub(j);
                        if (j == 3) v = v + ub(0) + ub(1) + ub(2);
                        if (j == dn - 4) v = v + ub(dn - 3) + ub(dn - 2) + ub(dn - 1);
                        return v;
                    };
                    T dtu = T(0);
                    for (int mm = -3; mm <= 3; ++mm) dtu = dtu + st.w1(mm) * U(rb - mm);
                    Py[nx_][col + (size_t)rb * NX] = A * Py[cur][col + (size_t)rb * NX] + dtu;
                    Xy[nx_][col + (size_t)rb * NX] = g.c2[i] * an[i] + A * Xy[cur][col + (size_t)rb * NX];
                }
            }
        }
        cur = nx_;
        if (grad_vp && hist) {
            const T* hk = hist + (size_t)k * n;
            const T* h1 = hist + (size_t)(k - 1) * n;
            const T* h2 = hist + (size_t)(k - 2) * n;
            for (size_t i = 0; i < n; ++i)
                gc2[i] += (double)an[i] * (((double)hk[i] - 2.0 * (double)h1[i] + (double)h2[i]) / (double)g.c2[i]);
        }
        if (srcadj)
            for (int j = 0; j < ns; ++j)
                if (acq.src_of_cell_tr[j] >= 0)
                    srcadj[(size_t)j * nt + (k - 2)] = g.c2[acq.src_cell[j]] * an[acq.src_cell[j]];
        std::swap(a2, a1);
        std::swap(a1, an);
    }
    if (grad_vp) {
        std::fill(grad_vp, grad_vp + (size_t)g.ny * g.nx, 0.0);
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) {
                const int jm = std::min(std::max(y - dn, 0), g.ny - 1), im = std::min(std::max(x - dn, 0), g.nx - 1);
                const double v = vp[(size_t)im * g.ny + jm];
                grad_vp[(size_t)im * g.ny + jm] += gc2[g.at(y, x)] * (2.0 * v * (double)dt * (double)dt) /
                                                  ((double)d * (double)d);
            }
    }
}

// ---------------------------------------------------------------------------
// Akima interpolation, ACM Algorithm 433 as modified in akima433.f90:6-170
// (INTRPL).  The Fortran reuses the coefficients of the previous point when
// it falls in the same interval (label 30) and shares storage between dead
// temporaries through EQUIVALENCE; both leave the values unchanged, so each
// point is evaluated on its own here.  Returns the routine's ERR code.
template <class T>
int intrpl(int L, const T* X, const T* Y, int N, const T* U, T* V) {
    if (L - 2 < 0) return 1;
    if (N <= 0) return 2;
    for (int i = 1; i < L; ++i) {
        if (X[i - 1] - X[i] < T(0)) continue;
        return X[i - 1] - X[i] == T(0) ? 6 : 7;
    }
    const int L0 = L, LM1 = L0 - 1, LM2 = LM1 - 1, LP1 = L0 + 1;
    auto x = [&](int i) { return X[i - 1]; };   // 1-based
    auto y = [&](int i) { return Y[i - 1]; };
    for (int k = 0; k < N; ++k) {
        const T UK = U[k];
        int I;
        if (LM2 == 0) I = 2;
        else if (UK >= x(L0)) I = LP1;
        else if (UK < x(1)) I = 1;
        else {
            int IMN = 2, IMX = L0;
            do {
                I = (IMN + IMX) / 2;
                if (UK >= x(I)) IMN = I + 1;
                else IMX = I;
            } while (IMX > IMN);
            I = IMX;
        }
        int J = I;
        if (J == 1) J = 2;
        if (J == LP1) J = L0;
        T X3 = x(J - 1), Y3 = y(J - 1), X4 = x(J), Y4 = y(J);
        T A3 = X4 - X3;
        T M3 = (Y4 - Y3) / A3;
        T M1 = 0, M2 = 0, M4 = 0, M5 = 0, A2 = 0, A4 = 0, X2 = 0, Y2 = 0, X5 = 0, Y5 = 0;
        if (LM2 == 0) {
            M2 = M3;
            M4 = M3;
        } else {
            bool to41 = J == 2;
            if (!to41) {
                X2 = x(J - 2); Y2 = y(J - 2);
                A2 = X3 - X2;
                M2 = (Y3 - Y2) / A2;
                if (J == L0) M4 = M3 + M3 - M2;
                else to41 = true;
            }
            if (to41) {
                X5 = x(J + 1); Y5 = y(J + 1);
                A4 = X5 - X4;
                M4 = (Y5 - Y4) / A4;
                if (J == 2) M2 = M3 + M3 - M4;
            }
        }
        if (J <= 3) M1 = M2 + M2 - M3;
        else {
            const T A1 = X2 - x(J - 3);
            M1 = (Y2 - y(J - 3)) / A1;
        }
        if (J >= LM1) M5 = M4 + M4 - M3;
        else {
            const T A5 = x(J + 2) - X5;
            M5 = (y(J + 2) - Y5) / A5;
        }
        T T3 = 0, T4 = 0;
        bool lp1 = I == LP1;
        if (!lp1) {
            T W2 = std::fabs(M4 - M3), W3 = std::fabs(M2 - M1);
            T SW = W2 + W3;
            if (!(SW != T(0))) { W2 = T(0.5); W3 = T(0.5); SW = T(1); }
            T3 = (W2 * M2 + W3 * M3) / SW;
        }
        if (I == 1) {
            T4 = T3;
            const T SA = A3 + A4;
            T3 = T(0.5) * (M1 + M2 - A4 * (A3 - A4) * (M3 - M4) / (SA * SA));
            X3 = X3 - A4;
            Y3 = Y3 - M2 * A4;
            A3 = A4;
            M3 = M2;
        } else {
            T W3 = std::fabs(M5 - M4), W4 = std::fabs(M3 - M2);
            T SW = W3 + W4;
            if (!(SW != T(0))) { W3 = T(0.5); W4 = T(0.5); SW = T(1); }
            T4 = (W3 * M3 + W4 * M4) / SW;
            if (lp1) {
                T3 = T4;
                const T SA = A2 + A3;
                T4 = T(0.5) * (M4 + M5 - A2 * (A2 - A3) * (M2 - M3) / (SA * SA));
                X3 = X4;
                Y3 = Y4;
                A3 = A2;
                M3 = M4;
            }
        }
        const T Q2 = (T(2) * (M3 - T3) + M3 - T4) / A3;
        const T Q3 = (-M3 - M3 + T3 + T4) / (A3 * A3);
        const T DX = UK - X3;
        V[k] = Y3 + DX * (T3 + DX * (Q2 + DX * Q3));
    }
    return 0;
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

// Exact discrete adjoint pair (see exact_forward / exact_adjoint).
#define ORACLE_EXACT(T, SUF)                                                                                \
    int oracle_exact_forward_##SUF(const oracle_opts* o, int ns, const int* nxs, const int* nys, int nr,   \
                                   const int* nxr, const int* nyr, const T* S, T d, T dt, int ny, int nx,  \
                                   const T* vp, int nt, T* rec, T* hist) {                                  \
        Opts op{o->dpml, o->free_surface};                                                                  \
        if (op.dpml < 2 || nt < 2) return -1;                                                               \
        ExactGrid<T> g(op, ny, nx, vp, d, dt);                                                              \
        exact_forward<T>(g, exact_acq<T>(g, ns, nxs, nys, nr, nxr, nyr), nt, S, rec, hist);                \
        return 0;                                                                                           \
    }                                                                                                       \
    int oracle_exact_adjoint_##SUF(const oracle_opts* o, int ns, const int* nxs, const int* nys, int nr,   \
                                   const int* nxr, const int* nyr, T d, T dt, int ny, int nx, const T* vp, \
                                   int nt, const T* e, const T* hist, T* srcadj, double* grad_vp) {         \
        Opts op{o->dpml, o->free_surface};                                                                  \
        if (op.dpml < 2 || nt < 2) return -1;                                                               \
        ExactGrid<T> g(op, ny, nx, vp, d, dt);                                                              \
        exact_adjoint<T>(g, exact_acq<T>(g, ns, nxs, nys, nr, nxr, nyr), nt, e, hist, d, dt, vp, srcadj,   \
                         grad_vp);                                                                          \
        return 0;                                                                                           \
    }

ORACLE_EXACT(float, f32)
ORACLE_EXACT(double, f64)

// akima433.f90 INTRPL on arbitrary increasing abscissae.
int oracle_intrpl_f32(int L, const float* X, const float* Y, int N, const float* U, float* V) {
    return intrpl<float>(L, X, Y, N, U, V);
}

// interpolation_akima.f90:1-31: every trace of Data_tmp(nt_Data, NumRec) on
// the grid (i-1)*dt_Data resampled onto (i-1)*dt, nt samples.
int oracle_akima_f32(const float* in, float dt_in, int nt_in, float* out, int ntr, float dt_out, int nt_out) {
    std::vector<float> X(nt_in), U(nt_out);
    for (int i = 0; i < nt_in; ++i) X[i] = float(i) * dt_in;
    for (int i = 0; i < nt_out; ++i) U[i] = float(i) * dt_out;
    int err = 0;
    for (int r = 0; r < ntr; ++r) {
        const int e = intrpl<float>(nt_in, X.data(), in + (size_t)r * nt_in, nt_out, U.data(), out + (size_t)r * nt_out);
        if (e) err = e;
    }
    return err;
}

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

/*
 * ORACLE — test infrastructure only. CPU FP64 restatement of the reference's
 * far-field CUR path, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. Nothing in the product links,
 * imports or calls this file.
 *
 * Follows (paths under the reference repository):
 *   element      codes/GPU_implementation/seed/user_impedance_for_julia_cuda.cu:178-214 (declaration
 *                only, body binary-only; restated from the textbook EFIE/RWG Galerkin form with
 *                the 3-point rule of CU:34), RWG layout get_edge.m:96-121
 *   pinv         CU:1873-1924, kernel_Sinv CU:1435-1444; my_pinv/invert_svec
 *                compres_single_block_spheres_for_sweep2.jl:1001-1022
 *   residual     CU:1928-1966
 *   sampling     compres_single_block_spheres_for_sweep2.jl:492-497 (uniform without
 *                replacement, sorted), drawn from Philox4x32-10 + selection sampling
 *   adaptive     cuda_cur_compress3, compres_single_block_spheres_for_sweep2.jl:878-973
 *   pivot rows   cur_smart_pinv2, compres_single_block_spheres.jl:2870-2887
 *
 * Pins (tests/): element vs a high-order quadrature of the explicit RWG
 * integrals and vs the closed-form Hertzian-dipole field; pinv vs LAPACK SVD;
 * residual vs numpy; Philox vs the published known-answer vectors; sampling
 * invariants; adaptive loop on synthetic exact-rank matrices. Element-wise
 * agreement with the reference's own (binary-only) element is not pinned.
 *
 * Geometry arguments: vertex 3 x nv (xyz per vertex), topol 3 x nt and edges
 * 4 x ne as 0-based int32, ln ne.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    const double* vertex;
    const int32_t* topol;
    const int32_t* edges;
    const double* ln;
    int ne;
    double k, eta;
} orc_geom;

/* ------------------------------------------------------------------ element */

static void tri_points(const orc_geom* g, int t, double q[3][3], double* area) {
    const double* v[3];
    for (int a = 0; a < 3; ++a) v[a] = &g->vertex[3 * g->topol[3 * t + a]];
    /* Degree-2 three-point rule: barycentric (2/3, 1/6, 1/6) and permutations. */
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int a = 0; a < 3; ++a) s += (a == i ? 2.0 / 3.0 : 1.0 / 6.0) * v[a][c];
            q[i][c] = s;
        }
    double e1[3], e2[3], cr[3];
    for (int c = 0; c < 3; ++c) {
        e1[c] = v[1][c] - v[0][c];
        e2[c] = v[2][c] - v[0][c];
    }
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    *area = 0.5 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
}

/*
 * Z_mn = -j k eta  sum_{T in m} sum_{T' in n}  A A' sum_{i,j} w_i w_j
 *        [ f_m(q_i) . f_n(q'_j) - div f_m div' f_n / k^2 ] G(|q_i - q'_j|),
 * G(R) = exp(-j k R) / (4 pi R), f = s l/(2A) (r - v_free), div f = s l / A.
 */
void orc_element(const orc_geom* g, int m, int n, double* zre, double* zim) {
    const double k = g->k;
    double sr = 0, si = 0;
    for (int hm = 0; hm < 2; ++hm) {
        const int tm = g->edges[4 * m + hm];
        const double* vm = &g->vertex[3 * g->edges[4 * m + 2 + hm]];
        const double sm = hm == 0 ? 1.0 : -1.0;
        double qm[3][3], am;
        tri_points(g, tm, qm, &am);
        const double fm = sm * g->ln[m] / (2 * am), dm = sm * g->ln[m] / am;
        for (int hn = 0; hn < 2; ++hn) {
            const int tn = g->edges[4 * n + hn];
            const double* vn = &g->vertex[3 * g->edges[4 * n + 2 + hn]];
            const double sn = hn == 0 ? 1.0 : -1.0;
            double qn[3][3], an;
            tri_points(g, tn, qn, &an);
            const double fn = sn * g->ln[n] / (2 * an), dn = sn * g->ln[n] / an;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double ff = 0, r2 = 0;
                    for (int c = 0; c < 3; ++c) {
                        ff += fm * (qm[i][c] - vm[c]) * fn * (qn[j][c] - vn[c]);
                        const double d = qm[i][c] - qn[j][c];
                        r2 += d * d;
                    }
                    const double R = sqrt(r2);
                    const double br = (ff - dm * dn / (k * k)) * am * an / 9.0;
                    const double gr = cos(k * R) / (4 * M_PI * R), gi = -sin(k * R) / (4 * M_PI * R);
                    sr += br * gr;
                    si += br * gi;
                }
        }
    }
    /* multiply by -j k eta: (-j)(sr + j si) = si - j sr */
    *zre = g->k * g->eta * si;
    *zim = -g->k * g->eta * sr;
}

/* Z[i + j*nr] = element(rows[i], cols[j]) (interleaved complex). */
void orc_fill(const orc_geom* g, const int32_t* rows, int nr, const int32_t* cols, int nc, double* Z) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < nc; ++j)
        for (int i = 0; i < nr; ++i) {
            double re, im;
            orc_element(g, rows[i], cols[j], &re, &im);
            Z[2 * ((size_t)j * nr + i)] = re;
            Z[2 * ((size_t)j * nr + i) + 1] = im;
        }
}

/* y = Z v without storing Z (rows x cols block). */
void orc_block_matvec(const orc_geom* g, const int32_t* rows, int nr, const int32_t* cols, int nc, const double* v,
                      double* y) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < nr; ++i) {
        double ar = 0, ai = 0;
        for (int j = 0; j < nc; ++j) {
            double re, im;
            orc_element(g, rows[i], cols[j], &re, &im);
            ar += re * v[2 * j] - im * v[2 * j + 1];
            ai += re * v[2 * j + 1] + im * v[2 * j];
        }
        y[2 * i] = ar;
        y[2 * i + 1] = ai;
    }
}

/* ------------------------------------------------------------------ near field
 * Near-field element: the reference integrates triangle pairs closer than
 * its Rinteg_s / Ranal_s radii with singular rules inside its binary-only
 * element (options at compres_single_block_spheres_for_sweep2.jl:1095-1098,
 * declaration user_impedance_for_julia_cuda.cu:178-214). Restated here with
 * the textbook singularity extraction: G = 1/(4 pi R) + (e^{-jkR} - 1)/(4 pi R),
 * the 1/R part integrated over the source triangle in closed form (Wilton,
 * Rao, Glisson, Schaubert, Al-Bundak, Butler, "Potential integrals for
 * uniform and linear source distributions on polygonal and polyhedral
 * domains", IEEE Trans. AP-32(3), 1984), the bounded remainder and the test
 * triangle by the 7-point degree-5 rule (Dunavant 1985). A pair is near when
 * the centroid distance is below near_factor x the longest edge of the two
 * triangles; other pairs keep the 3 x 3 rule of orc_element.
 * Pins (tests/test_oracle.py): the closed forms against polar quadrature
 * around the projected point (in plane, above, on a vertex's extension) and
 * against brute force away from the triangle; the element against the far
 * rule for separated pairs; reciprocity; the Mie series of a PEC sphere.
 */
static const double kD7[7][4] = {
    /* barycentric a, b, c and weight (sum 1) */
    {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0.225},
    {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506},
    {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.125939180544827},
};

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Is = int_T 1/|r - r'| dS', Iv = int_T (r' - rho)/|r - r'| dS' with rho the
 * projection of r on the plane of T (v0, v1, v2 counterclockwise about the
 * normal n = (v1 - v0) x (v2 - v0) / |...|). Wilton et al. 1984, eqs. for
 * uniform and linear sources: per edge i (a -> b), unit tangent l, outward
 * in-plane normal u = l x n, l+- = (b|a - rho).l, P0 = (a - rho).u,
 * R0^2 = P0^2 + h^2, R+- = |r - b|a|, f = ln((R+ + l+)/(R- + l-)):
 *   Is = sum P0 f - |h| sum [atan(P0 l+ / (R0^2 + |h| R+)) - atan(P0 l- / (R0^2 + |h| R-))]
 *   Iv = 1/2 sum u [R0^2 f + l+ R+ - l- R-]. */
void orc_potentials(const double* v0, const double* v1, const double* v2, const double* r, double* Is, double* Iv) {
    const double* V[3] = {v0, v1, v2};
    double e1[3], e2[3], n[3];
    for (int c = 0; c < 3; ++c) {
        e1[c] = v1[c] - v0[c];
        e2[c] = v2[c] - v0[c];
    }
    cross3(e1, e2, n);
    const double nn = sqrt(dot3(n, n));
    for (int c = 0; c < 3; ++c) n[c] /= nn;
    double d0[3];
    for (int c = 0; c < 3; ++c) d0[c] = r[c] - v0[c];
    const double h = dot3(d0, n), ah = fabs(h);
    double rho[3];
    for (int c = 0; c < 3; ++c) rho[c] = r[c] - h * n[c];
    double is = 0, iv[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const double* a = V[i];
        const double* b = V[(i + 1) % 3];
        double l[3], u[3], am[3], bm[3];
        for (int c = 0; c < 3; ++c) {
            l[c] = b[c] - a[c];
            am[c] = a[c] - rho[c];
            bm[c] = b[c] - rho[c];
        }
        const double len = sqrt(dot3(l, l));
        for (int c = 0; c < 3; ++c) l[c] /= len;
        cross3(l, n, u);
        const double lp = dot3(bm, l), lm = dot3(am, l), p0 = dot3(am, u);
        const double r02 = p0 * p0 + h * h;
        double ra[3], rb[3];
        for (int c = 0; c < 3; ++c) {
            ra[c] = r[c] - a[c];
            rb[c] = r[c] - b[c];
        }
        const double rp = sqrt(dot3(rb, rb)), rm = sqrt(dot3(ra, ra));
        double f = 0;
        if (r02 > 1e-28 * len * len) {
            /* the form without cancellation: (R + l)(R - l) = R0^2 */
            f = (lp + lm >= 0) ? log((rp + lp) / (rm + lm)) : log((rm - lm) / (rp - lp));
        } else if (lm > 0 || lp < 0) {
            /* on the edge's line, outside the segment: f finite, its weights P0 and R0^2 vanish */
            f = 0;
        }
        is += p0 * f;
        if (ah > 0) is -= ah * (atan2(p0 * lp, r02 + ah * rp) - atan2(p0 * lm, r02 + ah * rm));
        const double vv = 0.5 * (r02 * f + lp * rp - lm * rm);
        for (int c = 0; c < 3; ++c) iv[c] += vv * u[c];
    }
    *Is = is;
    for (int c = 0; c < 3; ++c) Iv[c] = iv[c];
}

static void tri_vertices(const orc_geom* g, int t, const double* v[3]) {
    for (int a = 0; a < 3; ++a) v[a] = &g->vertex[3 * g->topol[3 * t + a]];
}

static double tri_longest_edge(const double* v[3]) {
    double m = 0;
    for (int i = 0; i < 3; ++i) {
        double d[3];
        for (int c = 0; c < 3; ++c) d[c] = v[(i + 1) % 3][c] - v[i][c];
        m = fmax(m, sqrt(dot3(d, d)));
    }
    return m;
}

/* Singular-extraction integral of one triangle pair (test ta with free
 * vertex va, source tb with free vertex vb):
 *   J = int_Ta int_Tb [(r - va).(r' - vb) - 4/k^2] G dS' dS  (complex). */
static void near_pair(const orc_geom* g, int ta, const double* va, int tb, const double* vb, double* jr, double* ji) {
    const double k = g->k;
    const double *A[3], *B[3];
    tri_vertices(g, ta, A);
    tri_vertices(g, tb, B);
    double aa, ab, qa[3][3], qb[3][3];
    tri_points(g, ta, qa, &aa);
    tri_points(g, tb, qb, &ab);
    (void)qa;
    (void)qb;
    double nb[3], e1[3], e2[3];
    for (int c = 0; c < 3; ++c) {
        e1[c] = B[1][c] - B[0][c];
        e2[c] = B[2][c] - B[0][c];
    }
    cross3(e1, e2, nb);
    const double nn = sqrt(dot3(nb, nb));
    for (int c = 0; c < 3; ++c) nb[c] /= nn;
    double sr = 0, si = 0;
    for (int p = 0; p < 7; ++p) {
        double r[3];
        for (int c = 0; c < 3; ++c) r[c] = kD7[p][0] * A[0][c] + kD7[p][1] * A[1][c] + kD7[p][2] * A[2][c];
        const double wp = kD7[p][3] * aa;
        /* 1/R part in closed form */
        double is, iv[3];
        orc_potentials(B[0], B[1], B[2], r, &is, iv);
        const double h = (r[0] - B[0][0]) * nb[0] + (r[1] - B[0][1]) * nb[1] + (r[2] - B[0][2]) * nb[2];
        double i1s[3];  /* int (r' - vb)/R = Iv + (rho - vb) Is */
        for (int c = 0; c < 3; ++c) i1s[c] = iv[c] + (r[c] - h * nb[c] - vb[c]) * is;
        /* bounded remainder (e^{-jkR} - 1)/R by the 7-point rule */
        double i0r = 0, i0i = 0, i1r[3] = {0, 0, 0}, i1i[3] = {0, 0, 0};
        for (int q = 0; q < 7; ++q) {
            double rs[3], d[3];
            for (int c = 0; c < 3; ++c) {
                rs[c] = kD7[q][0] * B[0][c] + kD7[q][1] * B[1][c] + kD7[q][2] * B[2][c];
                d[c] = r[c] - rs[c];
            }
            const double R = sqrt(dot3(d, d));
            double gr, gi;  /* (e^{-jkR} - 1)/R, its limit -jk at R = 0 */
            if (R > 1e-12 * sqrt(ab)) {
                gr = (cos(k * R) - 1.0) / R;
                gi = -sin(k * R) / R;
            } else {
                gr = 0;
                gi = -k;
            }
            const double w = kD7[q][3] * ab;
            i0r += w * gr;
            i0i += w * gi;
            for (int c = 0; c < 3; ++c) {
                i1r[c] += w * gr * (rs[c] - vb[c]);
                i1i[c] += w * gi * (rs[c] - vb[c]);
            }
        }
        /* integrand (r - va).I1 - 4/k^2 I0, I0 = is + i0, I1 = i1s + i1 (times 1/4pi below) */
        double tr = 0, ti = 0;
        for (int c = 0; c < 3; ++c) {
            const double x = r[c] - va[c];
            tr += x * (i1s[c] + i1r[c]);
            ti += x * i1i[c];
        }
        const double c4 = 4.0 / (k * k);
        tr -= c4 * (is + i0r);
        ti -= c4 * i0i;
        sr += wp * tr;
        si += wp * ti;
    }
    *jr = sr / (4 * M_PI);
    *ji = si / (4 * M_PI);
}

/* Element with the near-pair rule (see above). */
void orc_element_near(const orc_geom* g, int m, int n, double near_factor, double* zre, double* zim) {
    const double k = g->k;
    double sr = 0, si = 0;  /* sum of s_a s_b l_m l_n / (4 A_a A_b) J */
    for (int hm = 0; hm < 2; ++hm) {
        const int tm = g->edges[4 * m + hm];
        const double* vm = &g->vertex[3 * g->edges[4 * m + 2 + hm]];
        const double sm = hm == 0 ? 1.0 : -1.0;
        double qm[3][3], am;
        tri_points(g, tm, qm, &am);
        const double *Vm[3];
        tri_vertices(g, tm, Vm);
        double cm[3];
        for (int c = 0; c < 3; ++c) cm[c] = (Vm[0][c] + Vm[1][c] + Vm[2][c]) / 3.0;
        for (int hn = 0; hn < 2; ++hn) {
            const int tn = g->edges[4 * n + hn];
            const double* vn = &g->vertex[3 * g->edges[4 * n + 2 + hn]];
            const double sn = hn == 0 ? 1.0 : -1.0;
            double qn[3][3], an;
            tri_points(g, tn, qn, &an);
            const double *Vn[3];
            tri_vertices(g, tn, Vn);
            double d[3];
            for (int c = 0; c < 3; ++c) d[c] = cm[c] - (Vn[0][c] + Vn[1][c] + Vn[2][c]) / 3.0;
            const double lim = near_factor * fmax(tri_longest_edge(Vm), tri_longest_edge(Vn));
            const double coef = sm * sn * g->ln[m] * g->ln[n] / (4 * am * an);
            if (dot3(d, d) < lim * lim) {
                double jr, ji;
                near_pair(g, tm, vm, tn, vn, &jr, &ji);
                sr += coef * jr;
                si += coef * ji;
            } else {
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        double pp = 0, r2 = 0;
                        for (int c = 0; c < 3; ++c) {
                            pp += (qm[i][c] - vm[c]) * (qn[j][c] - vn[c]);
                            const double e = qm[i][c] - qn[j][c];
                            r2 += e * e;
                        }
                        const double R = sqrt(r2);
                        const double br = (pp - 4.0 / (k * k)) * am * an / 9.0;
                        sr += coef * br * cos(k * R) / (4 * M_PI * R);
                        si -= coef * br * sin(k * R) / (4 * M_PI * R);
                    }
            }
        }
    }
    *zre = k * g->eta * si;
    *zim = -k * g->eta * sr;
}

void orc_fill_near(const orc_geom* g, const int32_t* rows, int nr, const int32_t* cols, int nc, double near_factor,
                   double* Z) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < nc; ++j)
        for (int i = 0; i < nr; ++i) {
            double re, im;
            orc_element_near(g, rows[i], cols[j], near_factor, &re, &im);
            Z[2 * ((size_t)j * nr + i)] = re;
            Z[2 * ((size_t)j * nr + i) + 1] = im;
        }
}

/* ------------------------------------------------------------------ pinv */
/*
 * One-sided Jacobi SVD of the r x r complex matrix W (column-major,
 * interleaved), cyclic-by-row sweeps, then
 * U = sum_j [s_j >= tau] v_j a_j^H / s_j^2,  tau = max(atol, rtol * s_max).
 * sigma (r) receives singular values (column order), returns the kept count.
 */
int orc_pinv(const double* W, int r, double atol, double rtol, double* U, double* sigma) {
    double* A = (double*)malloc(sizeof(double) * 2 * r * r);
    double* V = (double*)calloc(2 * (size_t)r * r, sizeof(double));
    memcpy(A, W, sizeof(double) * 2 * r * r);
    for (int i = 0; i < r; ++i) V[2 * ((size_t)i * r + i)] = 1.0;
    const double tol = fmax(1e-14, 8.8e-16 * r);
    for (int sweep = 0; sweep < 60; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < r - 1; ++p)
            for (int q = p + 1; q < r; ++q) {
                double* ap = A + 2 * (size_t)p * r;
                double* aq = A + 2 * (size_t)q * r;
                double al = 0, be = 0, gr = 0, gi = 0;
                for (int i = 0; i < r; ++i) {
                    const double a0 = ap[2 * i], a1 = ap[2 * i + 1], b0 = aq[2 * i], b1 = aq[2 * i + 1];
                    al += a0 * a0 + a1 * a1;
                    be += b0 * b0 + b1 * b1;
                    gr += a0 * b0 + a1 * b1; /* conj(a) b */
                    gi += a0 * b1 - a1 * b0;
                }
                const double gabs = hypot(gr, gi);
                if (!(gabs > tol * sqrt(al * be))) continue;
                rotated = 1;
                const double zeta = (be - al) / (2 * gabs);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                const double c = 1 / sqrt(1 + t * t), s = c * t;
                const double er = gr / gabs, ei = gi / gabs;
                for (int pass = 0; pass < 2; ++pass) {
                    double* xp = pass == 0 ? ap : V + 2 * (size_t)p * r;
                    double* xq = pass == 0 ? aq : V + 2 * (size_t)q * r;
                    for (int i = 0; i < r; ++i) {
                        const double a0 = xp[2 * i], a1 = xp[2 * i + 1];
                        /* b conj(e) */
                        const double b0 = xq[2 * i] * er + xq[2 * i + 1] * ei;
                        const double b1 = xq[2 * i + 1] * er - xq[2 * i] * ei;
                        xp[2 * i] = c * a0 - s * b0;
                        xp[2 * i + 1] = c * a1 - s * b1;
                        xq[2 * i] = s * a0 + c * b0;
                        xq[2 * i + 1] = s * a1 + c * b1;
                    }
                }
            }
        if (!rotated) break;
    }
    double smax = 0;
    double* w = (double*)malloc(sizeof(double) * r);
    for (int j = 0; j < r; ++j) {
        double nn = 0;
        for (int i = 0; i < r; ++i) {
            const double a0 = A[2 * ((size_t)j * r + i)], a1 = A[2 * ((size_t)j * r + i) + 1];
            nn += a0 * a0 + a1 * a1;
        }
        sigma[j] = sqrt(nn);
        w[j] = nn;
        if (sigma[j] > smax) smax = sigma[j];
    }
    const double tau = fmax(atol, rtol * smax);
    int kept = 0;
    for (int j = 0; j < r; ++j) {
        if (sigma[j] >= tau && w[j] > 0) {
            w[j] = 1.0 / w[j];
            ++kept;
        } else {
            w[j] = 0;
        }
    }
    for (int kk = 0; kk < r; ++kk)
        for (int i = 0; i < r; ++i) {
            double ur = 0, ui = 0;
            for (int j = 0; j < r; ++j) {
                if (w[j] == 0) continue;
                const double v0 = V[2 * ((size_t)j * r + i)], v1 = V[2 * ((size_t)j * r + i) + 1];
                const double a0 = A[2 * ((size_t)j * r + kk)], a1 = A[2 * ((size_t)j * r + kk) + 1];
                ur += w[j] * (v0 * a0 + v1 * a1);
                ui += w[j] * (v1 * a0 - v0 * a1);
            }
            U[2 * ((size_t)kk * r + i)] = ur;
            U[2 * ((size_t)kk * r + i) + 1] = ui;
        }
    free(w);
    free(A);
    free(V);
    return kept;
}

/* ------------------------------------------------------------------ residual */
static void cgemv(const double* A, int M, int N, const double* x, double* y) {
    for (int i = 0; i < M; ++i) {
        y[2 * i] = 0;
        y[2 * i + 1] = 0;
    }
    for (int j = 0; j < N; ++j) {
        const double xr = x[2 * j], xi = x[2 * j + 1];
        const double* a = A + 2 * (size_t)j * M;
        for (int i = 0; i < M; ++i) {
            y[2 * i] += a[2 * i] * xr - a[2 * i + 1] * xi;
            y[2 * i + 1] += a[2 * i] * xi + a[2 * i + 1] * xr;
        }
    }
}

/* y_new = C (U (R v)); returns ||y_old - y_new|| / ||y_new|| (CU:1928-1966). */
double orc_residual(const double* C, int m, int r, const double* U, const double* R, int n, const double* v,
                    const double* y_old, double* y_new) {
    double* t1 = (double*)malloc(sizeof(double) * 2 * r);
    double* t2 = (double*)malloc(sizeof(double) * 2 * r);
    cgemv(R, r, n, v, t1);
    cgemv(U, r, r, t1, t2);
    cgemv(C, m, r, t2, y_new);
    double nn = 0, dd = 0;
    for (int i = 0; i < m; ++i) {
        nn += y_new[2 * i] * y_new[2 * i] + y_new[2 * i + 1] * y_new[2 * i + 1];
        const double a = y_old[2 * i] - y_new[2 * i], b = y_old[2 * i + 1] - y_new[2 * i + 1];
        dd += a * a + b * b;
    }
    free(t1);
    free(t2);
    return sqrt(dd) / sqrt(nn);
}

/* ------------------------------------------------------------------ RNG + sampling */
/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x = ctr[0], y = ctr[1], z = ctr[2], w = ctr[3], k0 = key[0], k1 = key[1];
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * x, p1 = (uint64_t)0xCD9E8D57u * z;
        const uint32_t nx = (uint32_t)(p1 >> 32) ^ y ^ k0;
        const uint32_t ny = (uint32_t)p1;
        const uint32_t nz = (uint32_t)(p0 >> 32) ^ w ^ k1;
        const uint32_t nw = (uint32_t)p0;
        x = nx;
        y = ny;
        z = nz;
        w = nw;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
    out[3] = w;
}

static uint32_t stream_word(uint64_t seed, uint32_t block, uint32_t round, uint32_t purpose, uint32_t t) {
    const uint32_t ctr[4] = {t / 4, block, (round << 8) | purpose, 0x46435552u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox(ctr, key, o);
    return o[t % 4];
}

/* r of 0..N-1, sorted, uniform without replacement (selection sampling, Knuth
 * Algorithm S, with the integer-exact acceptance floor(u (N-t) / 2^32) < r - got). */
void orc_sample(uint64_t seed, int block, int round, int purpose, int N, int r, int32_t* out) {
    int got = 0;
    for (int t = 0; t < N && got < r; ++t) {
        const uint32_t u = stream_word(seed, (uint32_t)block, (uint32_t)round, (uint32_t)purpose, (uint32_t)t);
        if ((uint32_t)(((uint64_t)u * (uint32_t)(N - t)) >> 32) < (uint32_t)(r - got)) out[got++] = t;
    }
}

/* Learned mode: h held-out rows of 0..m-1 (stream purpose 4), then r sample
 * rows among the remaining m - h positions (purpose 2, word = candidate index). */
void orc_sample_heldout(uint64_t seed, int block, int round, int m, int h, int r, int32_t* hpos, int32_t* rows) {
    orc_sample(seed, block, round, 4, m, h, hpos);
    int got = 0, cand = 0, hk = 0;
    for (int t = 0; t < m && got < r; ++t) {
        if (hk < h && hpos[hk] == t) {
            ++hk;
            continue;
        }
        const uint32_t u = stream_word(seed, (uint32_t)block, (uint32_t)round, 2u, (uint32_t)cand);
        if ((uint32_t)(((uint64_t)u * (uint32_t)(m - h - cand)) >> 32) < (uint32_t)(r - got)) rows[got++] = t;
        ++cand;
    }
}

static double u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

void orc_vtest(uint64_t seed, int block, int n, double* v) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int j = 0; j < n; ++j) {
        const uint32_t ctr[4] = {(uint32_t)j, (uint32_t)block, 3u, 0x46435552u};
        uint32_t o[4];
        orc_philox(ctr, key, o);
        v[2 * j] = u01(o[0], o[1]);
        v[2 * j + 1] = u01(o[2], o[3]);
    }
}

/* rows[j] = argmax_i |C[i, j]| (first on ties), 0-based positions. */
void orc_pivot_rows(const double* C, int m, int r, int32_t* rows) {
    for (int j = 0; j < r; ++j) {
        double best = -1;
        int bi = 0;
        for (int i = 0; i < m; ++i) {
            const double a = C[2 * ((size_t)j * m + i)], b = C[2 * ((size_t)j * m + i) + 1];
            const double v = a * a + b * b;
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        rows[j] = bi;
    }
}

/* ------------------------------------------------------------------ adaptive loop */
typedef struct {
    double tol, pinv_tol, pinv_rtol;
    uint64_t seed;
    int row_mode; /* 0 uniform, 1 pivot */
    int r_start;  /* 0: D < 10 ? 1 : D / 10 */
    int max_rank; /* 0: D */
    int block_id;
} orc_opts;

typedef struct {
    int r, rounds, kept, converged;
    double err;
} orc_info;

/* Entry provider: EFIE geometry (g != NULL) or an explicit dense m x n matrix. */
typedef struct {
    const orc_geom* g;
    const int32_t* rows;
    const int32_t* cols;
    const double* dense; /* m x n interleaved, used when g == NULL */
    int m, n;
} orc_block;

static void block_panel(const orc_block* b, const int32_t* ri, int nr, const int32_t* ci, int nc, double* out) {
    /* ri, ci: positions within the block */
    if (b->g) {
        int32_t* rr = (int32_t*)malloc(sizeof(int32_t) * nr);
        int32_t* cc = (int32_t*)malloc(sizeof(int32_t) * nc);
        for (int i = 0; i < nr; ++i) rr[i] = b->rows[ri[i]];
        for (int j = 0; j < nc; ++j) cc[j] = b->cols[ci[j]];
        orc_fill(b->g, rr, nr, cc, nc, out);
        free(rr);
        free(cc);
    } else {
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nr; ++i) {
                out[2 * ((size_t)j * nr + i)] = b->dense[2 * ((size_t)ci[j] * b->m + ri[i])];
                out[2 * ((size_t)j * nr + i) + 1] = b->dense[2 * ((size_t)ci[j] * b->m + ri[i]) + 1];
            }
    }
}

/*
 * cuda_cur_compress3: r0 = D < 10 ? 1 : D/10; first round against ones(M)
 * (its error discarded), then r doubles (capped at D) until err <= tol or the
 * round ran at r = D. Final C (m x r), U (r x r), R (r x n) and samples
 * (positions in the block) go to caller buffers sized for max_rank (or D).
 */
int orc_compress_block(const orc_block* b, const orc_opts* o, double* Cout, double* Uout, double* Rout,
                       int32_t* rs_out, int32_t* cs_out, double* y_out, orc_info* info) {
    const int m = b->m, n = b->n, D = m < n ? m : n;
    const int cap = o->max_rank > 0 && o->max_rank < D ? o->max_rank : D;
    int r = o->r_start > 0 ? o->r_start : (D < 10 ? 1 : D / 10);
    if (r > cap) r = cap;
    double* v = (double*)malloc(sizeof(double) * 2 * n);
    double* y0 = (double*)malloc(sizeof(double) * 2 * m);
    double* y1 = (double*)malloc(sizeof(double) * 2 * m);
    orc_vtest(o->seed, o->block_id, n, v);
    for (int i = 0; i < m; ++i) {
        y0[2 * i] = 1;
        y0[2 * i + 1] = 0;
    }
    int32_t* all_rows = (int32_t*)malloc(sizeof(int32_t) * m);
    int32_t* all_cols = (int32_t*)malloc(sizeof(int32_t) * n);
    for (int i = 0; i < m; ++i) all_rows[i] = i;
    for (int j = 0; j < n; ++j) all_cols[j] = j;
    int round = 0, converged = 0, kept = 0;
    double err = INFINITY;
    int final_r = r;
    while (1) {
        int32_t* cs = (int32_t*)malloc(sizeof(int32_t) * r);
        int32_t* rs = (int32_t*)malloc(sizeof(int32_t) * r);
        double* C = (double*)malloc(sizeof(double) * 2 * (size_t)m * r);
        double* W = (double*)malloc(sizeof(double) * 2 * (size_t)r * r);
        double* U = (double*)malloc(sizeof(double) * 2 * (size_t)r * r);
        double* R = (double*)malloc(sizeof(double) * 2 * (size_t)r * n);
        double* sig = (double*)malloc(sizeof(double) * r);
        orc_sample(o->seed, o->block_id, round, 1, n, r, cs);
        block_panel(b, all_rows, m, cs, r, C);
        if (o->row_mode == 1)
            orc_pivot_rows(C, m, r, rs);
        else
            orc_sample(o->seed, o->block_id, round, 2, m, r, rs);
        block_panel(b, rs, r, cs, r, W);
        block_panel(b, rs, r, all_cols, n, R);
        kept = orc_pinv(W, r, o->pinv_tol, o->pinv_rtol, U, sig);
        const double e = orc_residual(C, m, r, U, R, n, v, y0, y1);
        double* t = y0;
        y0 = y1;
        y1 = t;
        final_r = r;
        memcpy(Cout, C, sizeof(double) * 2 * (size_t)m * r);
        memcpy(Uout, U, sizeof(double) * 2 * (size_t)r * r);
        memcpy(Rout, R, sizeof(double) * 2 * (size_t)r * n);
        memcpy(rs_out, rs, sizeof(int32_t) * r);
        memcpy(cs_out, cs, sizeof(int32_t) * r);
        free(cs);
        free(rs);
        free(C);
        free(W);
        free(U);
        free(R);
        free(sig);
        if (round == 0) {
            r = (2 * r < cap) ? 2 * r : cap;
            round = 1;
            continue;
        }
        err = e;
        ++round;
        if (!(err > o->tol)) {
            converged = err <= o->tol;
            break;
        }
        if (r >= cap) break;
        r = (2 * r <= cap) ? 2 * r : cap;
    }
    memcpy(y_out, y0, sizeof(double) * 2 * m);
    info->r = final_r;
    info->rounds = round;
    info->kept = kept;
    info->converged = converged;
    info->err = err;
    free(v);
    free(y0);
    free(y1);
    free(all_rows);
    free(all_cols);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Host-side species tables: transition list, state index and the Wigner-coupled
// SE coefficient term list.
//
// all_transitions (miscellaneous.f90:394-460), setup_count_array (206-239) and the
// constants of pop_eqs.f90:1-359 (L&L04 Eqs. 7.20a-f), which depend only on
// (J, K, J', K').  Wigner 3j/6j values are evaluated with Racah's single-sum
// formulas over exact integer factorial ratios held in long double; for the
// arguments used here (one j equal to 1 or K <= kmax) the sums have few terms and
// the result is correctly rounded to double in practice (pinned against exact
// values in the tests).
#include <cmath>
#include <cstdlib>
#include <map>
#include <vector>

#include "common.h"

namespace portal {

namespace {

long double fact(int n) {
    static std::vector<long double> t;
    if (t.empty()) {
        t.resize(400);
        t[0] = 1.0L;
        for (int i = 1; i < 400; i++) t[i] = t[i - 1] * i;
    }
    return t.at(n);
}

bool triangle(int a, int b, int c) { return !((a + b + c) & 1) && c <= a + b && c >= std::abs(a - b); }

long double delta(int a, int b, int c) {
    return fact((a + b - c) / 2) * fact((a - b + c) / 2) * fact((b + c - a) / 2) / fact((a + b + c) / 2 + 1);
}

}  // namespace

// Wigner 3j with doubled arguments (as WIGXJPF's fwig3jj).
double wig3jj(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || !triangle(j1, j2, j3)) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (((j1 + m1) | (j2 + m2) | (j3 + m3)) & 1) return 0.0;
    const int n1 = (j1 + j2 - j3) / 2, n2 = (j1 - m1) / 2, n3 = (j2 + m2) / 2;
    const int d1 = (j3 - j2 + m1) / 2, d2 = (j3 - j1 - m2) / 2;
    int lo = 0;
    if (-d1 > lo) lo = -d1;
    if (-d2 > lo) lo = -d2;
    int hi = n1;
    if (n2 < hi) hi = n2;
    if (n3 < hi) hi = n3;
    long double sum = 0.0L;
    for (int k = lo; k <= hi; k++) {
        const long double term =
            1.0L / (fact(k) * fact(d1 + k) * fact(d2 + k) * fact(n1 - k) * fact(n2 - k) * fact(n3 - k));
        sum += (k % 2) ? -term : term;
    }
    const long double norm = std::sqrt(delta(j1, j2, j3) * fact((j1 + m1) / 2) * fact((j1 - m1) / 2) *
                                       fact((j2 + m2) / 2) * fact((j2 - m2) / 2) * fact((j3 + m3) / 2) *
                                       fact((j3 - m3) / 2));
    long double r = norm * sum;
    if (((j1 - j2 - m3) / 2) % 2) r = -r;
    return (double)r;
}

// Wigner 6j with doubled arguments (as WIGXJPF's fwig6jj).
double wig6jj(int j1, int j2, int j3, int j4, int j5, int j6) {
    if (!triangle(j1, j2, j3) || !triangle(j1, j5, j6) || !triangle(j4, j2, j6) || !triangle(j4, j5, j3)) return 0.0;
    const int a[4] = {(j1 + j2 + j3) / 2, (j1 + j5 + j6) / 2, (j4 + j2 + j6) / 2, (j4 + j5 + j3) / 2};
    const int b[3] = {(j1 + j2 + j4 + j5) / 2, (j2 + j3 + j5 + j6) / 2, (j3 + j1 + j6 + j4) / 2};
    int lo = a[0], hi = b[0];
    for (int i = 1; i < 4; i++) lo = a[i] > lo ? a[i] : lo;
    for (int i = 1; i < 3; i++) hi = b[i] < hi ? b[i] : hi;
    long double sum = 0.0L;
    for (int t = lo; t <= hi; t++) {
        const long double term = fact(t + 1) / (fact(t - a[0]) * fact(t - a[1]) * fact(t - a[2]) * fact(t - a[3]) *
                                                fact(b[0] - t) * fact(b[1] - t) * fact(b[2] - t));
        sum += (t % 2) ? -term : term;
    }
    return (double)(std::sqrt(delta(j1, j2, j3) * delta(j1, j5, j6) * delta(j4, j2, j6) * delta(j4, j5, j3)) * sum);
}

namespace {

inline double w3(int a, int b, int c, int d, int e, int f) { return wig3jj(2 * a, 2 * b, 2 * c, 2 * d, 2 * e, 2 * f); }
inline double sgn(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }  // (-1)^n

// Sum over m, q of w1 w2 w3^2 w4 for the four "transfer/relaxation" constants.
// kind 0: t_A  (pop_eqs.f90:1-78)      level (j,k), lower (jl,kl)
// kind 1: t_S  (pop_eqs.f90:80-156)    level (j,k), upper (ju,ku)
// kind 2: r_A  (pop_eqs.f90:204-281)   level (j,k), upper ju, k'
// kind 3: r_S  (pop_eqs.f90:283-359)   level (j,k), lower jl, k'
void coupling(int kind, int j, int k, int j2, int k2, double &c0, double &c2) {
    c0 = c2 = 0.0;
    const int dj = j2 - j, dk = k2 - k;
    if (std::abs(dj) > 1 || !(dk == 0 || std::abs(dk) == 2)) return;
    const double f1 = std::sqrt(3.0 * (2.0 * k + 1.0) * (2.0 * k2 + 1.0));
    double s0 = 0.0, s2 = 0.0;
    for (int m = -j; m <= j; m++)
        for (int q = -1; q <= 1; q++) {
            double a = 0, b = 0, c = 0, sg = 1.0;
            switch (kind) {
                case 0: {
                    const int ml = q + m;
                    a = w3(j, j, k, m, -m, 0);
                    b = w3(j2, j2, k2, ml, -ml, 0);
                    c = w3(j, j2, 1, -m, ml, -q);
                    break;
                }
                case 1: {
                    const int mu = m - q;
                    a = w3(j, j, k, m, -m, 0);
                    b = w3(j2, j2, k2, mu, -mu, 0);
                    c = w3(j2, j, 1, -mu, m, -q);
                    break;
                }
                case 2: {
                    const int mu = m - q;
                    a = w3(j, j, k, m, -m, 0);
                    b = w3(j, j, k2, m, -m, 0);
                    c = w3(j2, j, 1, -mu, m, -q);
                    sg = sgn(1 + q);
                    break;
                }
                default: {
                    const int ml = m + q;
                    a = w3(j, j, k, m, -m, 0);
                    b = w3(j, j, k2, m, -m, 0);
                    c = w3(j, j2, 1, -m, ml, -q);
                    sg = sgn(1 - q);
                    break;
                }
            }
            const double ww = a * b * c * c;
            if (dk == 0) s0 += sg * ww * w3(1, 1, 0, q, -q, 0);
            s2 += sg * ww * w3(1, 1, 2, q, -q, 0);
        }
    if (kind <= 1) {
        const double fac = (2.0 * j2 + 1.0) * f1 * sgn(1 - dj);
        c0 = s0 * fac;
        c2 = s2 * std::sqrt(5.0) * fac;
    } else {
        const double fac = (2.0 * j + 1.0) * f1;
        c0 = s0 * fac;
        c2 = s2 * fac * std::sqrt(5.0);
    }
}

// t_E (pop_eqs.f90:158-201)
double t_spont(int j, int k, int ju, int ku) {
    if (std::abs(ju - j) > 1 || ku != k) return 0.0;
    return (2.0 * ju + 1.0) * sgn(1 + j + ju + k) * wig6jj(2 * ju, 2 * ju, 2 * k, 2 * j, 2 * j, 2);
}

}  // namespace

// coefficients exported for the tests (kind as in coupling(); kind 4 = t_E)
extern "C" void portal_debug_coupling(int kind, int j, int k, int j2, int k2, double *c0, double *c2) {
    if (kind == 4) {
        *c0 = t_spont(j, k, j2, k2);
        *c2 = 0.0;
    } else {
        coupling(kind, j, k, j2, k2, *c0, *c2);
    }
}
extern "C" double portal_debug_wig3jj(int a, int b, int c, int d, int e, int f) { return wig3jj(a, b, c, d, e, f); }
extern "C" double portal_debug_wig6jj(int a, int b, int c, int d, int e, int f) { return wig6jj(a, b, c, d, e, f); }

int build_species_tables(Species &s) {
    const int L = s.e_num;
    // all_transitions: unique (lower, upper) pairs ordered by lower then upper
    std::vector<std::vector<char>> T(L, std::vector<char>(L, 0));
    auto add = [&](const std::vector<int> &lst, int n) {
        for (int i = 0; i < n; i++) {
            const int j = lst[2 * i] - 1, k = lst[2 * i + 1] - 1;
            if (s.e_levs_hz[k] > s.e_levs_hz[j])
                T[j][k] = 1;
            else
                T[k][j] = 1;
        }
    };
    add(s.r_trans, s.r_num);
    add(s.c_trans, s.c_num);
    s.all_trans.clear();
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++)
            if (T[i][j]) {
                s.all_trans.push_back(i + 1);
                s.all_trans.push_back(j + 1);
            }
    s.n_trans = (int)s.all_trans.size() / 2;
    // setup_count_array
    s.count_array.assign(L, 0);
    int n = 0;
    for (int i = 0; i < L; i++) {
        s.count_array[i] = n;
        const int kt = std::min(2 * s.j_lev[i], s.kmax);
        for (int k = 0; k <= kt; k += 2) n++;
    }
    s.c_dim = n;
    s.trans_rad.assign(s.n_trans, -1);
    s.trans_col.assign(s.n_trans, -1);
    for (int t = 0; t < s.n_trans; t++) {
        const int lo = s.all_trans[2 * t], up = s.all_trans[2 * t + 1];
        for (int j = 0; j < s.r_num; j++)
            if (s.r_trans[2 * j] == lo && s.r_trans[2 * j + 1] == up) { s.trans_rad[t] = j; break; }
        for (int j = 0; j < s.c_num; j++)
            if (s.c_trans[2 * j] == lo && s.c_trans[2 * j + 1] == up) { s.trans_col[t] = j; break; }
    }

    // term list: entry -> [(scalar index, coefficient, class)]
    std::map<int, std::vector<std::pair<int, std::pair<double, int>>>> ent;
    const int cd = s.c_dim, kmax = s.kmax;
    auto kmaxof = [&](int lev) { return std::min(2 * s.j_lev[lev], kmax); };
    auto put = [&](int r, int c, int q, double coef, int cls) {
        if (coef != 0.0) ent[r + cd * c].push_back({q, {coef, cls}});
    };
    // scalars per transition; entries between K = 0 slots (the population block) take
    // J0 and A from their own copies, which the accelerated lambda iteration modifies
    enum { B21J0 = 0, B21J2, B12J0, B12J2, AA, C21, C12, B21J0P, B12J0P, AAP, NX };
    auto pop = [](int k1, int k2, int q) {
        if (k1 != 0 || k2 != 0) return q;
        return q == B21J0 ? (int)B21J0P : q == B12J0 ? (int)B12J0P : q == AA ? (int)AAP : q;
    };
    for (int t = 0; t < s.n_trans; t++) {
        const int lo = s.all_trans[2 * t] - 1, up = s.all_trans[2 * t + 1] - 1;
        const int jl = s.j_lev[lo], ju = s.j_lev[up];
        const double gl = 2.0 * jl + 1.0, gu = 2.0 * ju + 1.0;
        // transfer into the lower level from the upper one (pop_eqs.f90:467-515)
        for (int k1 = 0, c1 = s.count_array[lo]; k1 <= kmaxof(lo); k1 += 2, c1++)
            for (int k2 = 0, c2 = s.count_array[up]; k2 <= kmaxof(up); k2 += 2, c2++) {
                double ts0, ts2;
                coupling(1, jl, k1, ju, k2, ts0, ts2);
                put(c1, c2, NX * t + pop(k1, k2, B21J0), ts0, 1);
                put(c1, c2, NX * t + pop(k1, k2, B21J2), ts2, 1);
                put(c1, c2, NX * t + pop(k1, k2, AA), t_spont(jl, k1, ju, k2), 0);
                if (k1 == 0 && k2 == 0) put(c1, c2, NX * t + pop(k1, k2, C21), std::sqrt(gu / gl), 2);
            }
        // transfer into the upper level from the lower one (pop_eqs.f90:518-562)
        for (int k1 = 0, c1 = s.count_array[up]; k1 <= kmaxof(up); k1 += 2, c1++)
            for (int k2 = 0, c2 = s.count_array[lo]; k2 <= kmaxof(lo); k2 += 2, c2++) {
                double ta0, ta2;
                coupling(0, ju, k1, jl, k2, ta0, ta2);
                put(c1, c2, NX * t + pop(k1, k2, B12J0), ta0, 1);
                put(c1, c2, NX * t + pop(k1, k2, B12J2), ta2, 1);
                if (k1 == 0 && k2 == 0) put(c1, c2, NX * t + pop(k1, k2, C12), std::sqrt(gl / gu), 2);
            }
        // relaxation of the lower level: absorption and collisional excitation
        for (int k1 = 0, c1 = s.count_array[lo]; k1 <= kmaxof(lo); k1 += 2, c1++)
            for (int k2 = 0, c2 = s.count_array[lo]; k2 <= kmaxof(lo); k2 += 2, c2++) {
                double r0, r2;
                coupling(2, jl, k1, ju, k2, r0, r2);
                put(c1, c2, NX * t + pop(k1, k2, B12J0), -r0, 1);
                put(c1, c2, NX * t + pop(k1, k2, B12J2), -r2, 1);
                if (k1 == k2) put(c1, c2, NX * t + pop(k1, k2, C12), -1.0, 2);
            }
        // relaxation of the upper level: stimulated + spontaneous emission, de-excitation
        for (int k1 = 0, c1 = s.count_array[up]; k1 <= kmaxof(up); k1 += 2, c1++)
            for (int k2 = 0, c2 = s.count_array[up]; k2 <= kmaxof(up); k2 += 2, c2++) {
                double r0, r2;
                coupling(3, ju, k1, jl, k2, r0, r2);
                put(c1, c2, NX * t + pop(k1, k2, B21J0), -r0, 1);
                put(c1, c2, NX * t + pop(k1, k2, B21J2), -r2, 1);
                if (k1 == k2) {
                    put(c1, c2, NX * t + pop(k1, k2, AA), -1.0, 0);
                    put(c1, c2, NX * t + pop(k1, k2, C21), -1.0, 2);
                }
            }
    }
    SeTables &tb = s.se;
    tb.c_dim = cd;
    tb.n_trans = s.n_trans;
    tb.ent_idx.clear();
    tb.ent_ptr.assign(1, 0);
    tb.term_q.clear();
    tb.term_coef.clear();
    tb.term_class.clear();
    for (auto &kv : ent) {
        tb.ent_idx.push_back(kv.first);
        for (auto &tq : kv.second) {
            tb.term_q.push_back(tq.first);
            tb.term_coef.push_back(tq.second.first);
            tb.term_class.push_back(tq.second.second);
        }
        tb.ent_ptr.push_back((int)tb.term_q.size());
    }
    return 0;
}

// de_polarization (depol.f90:242-317): Morris (1985) Eq. B11 in the sudden
// approximation, the rank-K relaxation rate coefficient of a level with J,
//   D^(K)(J) = (2J+1)^2 sum_{m=1}^{2J} (-1)^m (2m+1) (J J m; 0 0 0)^2
//              [{J J m; J J 0} - {J J m; J J K}] k_m(T),
// for even K = 2..min(2J, kmax) at each tabulated collision temperature.  k_m is the
// tabulated rate coefficient between the J = 0 level and the level whose J is m (0.25 /
// 0.75 partner mix with two partners, as make_AC), 0 where the species has none.  The
// reference reads its rate array at level index m (J = m - 1); this takes J = m.  out
// is c_temp x e_num x (kmax - 1) column-major, the reference's D0(nT, nlev, 2:kmax).
// Returns 0, or -1 if the species has no J = 0 level.
int depol_table(const Species &s, std::vector<double> &out) {
    const int nT = s.c_temp, L = s.e_num, nk = std::max(s.kmax - 1, 0), nC = s.c_num;
    out.assign((size_t)nT * L * nk, 0.0);
    int l0 = -1;
    for (int i = 0; i < L && l0 < 0; i++)
        if (s.j_lev[i] == 0) l0 = i;
    if (l0 < 0) return -1;
    for (int i = 0; i < L; i++) {
        const int J = s.j_lev[i];
        const double lj2 = (2.0 * J + 1.0) * (2.0 * J + 1.0);
        for (int K = 2; K <= s.kmax && K <= 2 * J; K += 2)
            for (int m = 1; m <= 2 * J; m++) {
                const double w1 = wig3jj(2 * J, 2 * J, 2 * m, 0, 0, 0);
                if (w1 == 0.0) continue;
                int lm = -1, col = -1;
                for (int q = 0; q < L && lm < 0; q++)
                    if (q != l0 && s.j_lev[q] == m) lm = q;
                for (int c = 0; c < nC && lm >= 0 && col < 0; c++) {
                    const int a = s.c_trans[2 * c] - 1, b = s.c_trans[2 * c + 1] - 1;
                    if ((a == l0 && b == lm) || (a == lm && b == l0)) col = c;
                }
                if (col < 0) continue;
                const double f = lj2 * sgn(m) * (2.0 * m + 1.0) * w1 * w1 *
                                 (wig6jj(2 * J, 2 * J, 2 * m, 2 * J, 2 * J, 0) - wig6jj(2 * J, 2 * J, 2 * m, 2 * J, 2 * J, 2 * K));
                for (int t = 0; t < nT; t++) {
                    const double *C = s.t_Cji.data();
                    const double k = s.c_par == 2 ? 0.25 * C[t + (size_t)nT * col] + 0.75 * C[t + (size_t)nT * (col + nC)]
                                                  : C[t + (size_t)nT * col];
                    out[t + (size_t)nT * (i + (size_t)L * (K - 2))] += f * k;
                }
            }
    }
    return 0;
}

}  // namespace portal

// sbp_tables.h -- host-side construction of the dimensionless 1-D factors.
//
// From the reference's diagonal-norm SBP second-derivative tables
// (diagonal_sbp.jl:191-408, orders p = 2, 4, 6, 8, 10: bhinv, interior d,
// closure bd, boundary row BS)
// this builds, for a side of N points,
//     hw = H / h                        (H = diag(h / bhinv) near the faces, h inside)
//     m  = h (BS - H D2)                (the 1-D "stiffness", symmetric PSD, m 1 = 0)
//     k  = m - (E bs + bs^T E) + 13 E   (m plus the Dirichlet SAT with tau = -13/h, beta = 1)
// so that the reference's A = -H~(D2 + SAT) (level_2_multigrid_new.jl:141-185)
// equals  k (x) hw  +  hw (x) m  (first factor on the slow index j).
// The Neumann SAT with tau_N = -1 cancels the BS term of the fast index exactly.
// Only the first R rows / C columns differ from the interior stencil; they
// are extracted from a small dense build and checked for shape.  With the
// reference's tau = 13 the Dirichlet factor k is positive definite for
// p <= 6 but not for p = 8 (its smallest eigenvalue is about -2.7; the
// penalty would have to exceed about 800) nor p = 10 (about -7.8): those
// operators are symmetric indefinite, so their apply is exact but CG-type
// solves may break down.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

#include "sbp_kernels.h"

namespace sbp {

struct Sbp1D {
    int p = 0, bm = 0, bn = 0, nbs = 0;
    double bhinv[kMaxR] = {};
    double d[11] = {};                   // interior D2 stencil (times h^2), p + 1 entries
    double bd[kMaxR][kMaxC] = {};        // closure of D2 (times h^2), bm x bn
    double bs[kMaxR] = {};               // boundary derivative row (times h), nbs entries
};

// The reference's rationals (diagonal_sbp.jl:191-197, 198-208, 210-222, 224-311);
// for p = 10 its decimal M entries and the construction of bd (313-408).
inline bool sbp_1d(int p, Sbp1D* s) {
    *s = Sbp1D{};
    s->p = p;
    auto fill = [s](int bm, int bn, int nbs, const double* bh, const double* d, const double* bd, const double* bs) {
        s->bm = bm; s->bn = bn; s->nbs = nbs;
        for (int k = 0; k < bm; ++k) s->bhinv[k] = bh[k];
        for (int k = 0; k <= s->p; ++k) s->d[k] = d[k];
        for (int r = 0; r < bm; ++r)
            for (int c = 0; c < bn; ++c) s->bd[r][c] = bd[r * bn + c];
        for (int k = 0; k < nbs; ++k) s->bs[k] = bs[k];
    };
    if (p == 2) {
        const double bh[1] = {2.0};
        const double d[3] = {1.0, -2.0, 1.0};
        const double bd[1][3] = {{1.0, -2.0, 1.0}};
        const double bs[3] = {3.0 / 2.0, -2.0, 1.0 / 2.0};
        fill(1, 3, 3, bh, d, &bd[0][0], bs);
        return true;
    }
    if (p == 4) {
        const double bh[4] = {48.0 / 17.0, 48.0 / 59.0, 48.0 / 43.0, 48.0 / 49.0};
        const double d[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
        const double bd[4][6] = {
            {2.0, -5.0, 4.0, -1.0, 0.0, 0.0},
            {1.0, -2.0, 1.0, 0.0, 0.0, 0.0},
            {-4.0 / 43.0, 59.0 / 43.0, -110.0 / 43.0, 59.0 / 43.0, -4.0 / 43.0, 0.0},
            {-1.0 / 49.0, 0.0, 59.0 / 49.0, -118.0 / 49.0, 64.0 / 49.0, -4.0 / 49.0}};
        const double bs[4] = {11.0 / 6.0, -3.0, 3.0 / 2.0, -1.0 / 3.0};
        fill(4, 6, 4, bh, d, &bd[0][0], bs);
        return true;
    }
    if (p == 6) {
        const double bh[6] = {43200.0 / 13649.0, 8640.0 / 12013.0, 4320.0 / 2711.0, 4320.0 / 5359.0, 8640.0 / 7877.0, 43200.0 / 43801.0};
        const double d[7] = {1.0 / 90.0, -3.0 / 20.0, 3.0 / 2.0, -49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};
        const double bd[6][9] = {
            {114170.0 / 40947.0, -438107.0 / 54596.0, 336409.0 / 40947.0, -276997.0 / 81894.0, 3747.0 / 13649.0, 21035.0 / 163788.0, 0.0, 0.0, 0.0},
            {6173.0 / 5860.0, -2066.0 / 879.0, 3283.0 / 1758.0, -303.0 / 293.0, 2111.0 / 3516.0, -601.0 / 4395.0, 0.0, 0.0, 0.0},
            {-52391.0 / 81330.0, 134603.0 / 32532.0, -21982.0 / 2711.0, 112915.0 / 16266.0, -46969.0 / 16266.0, 30409.0 / 54220.0, 0.0, 0.0, 0.0},
            {68603.0 / 321540.0, -12423.0 / 10718.0, 112915.0 / 32154.0, -75934.0 / 16077.0, 53369.0 / 21436.0, -54899.0 / 160770.0, 48.0 / 5359.0, 0.0, 0.0},
            {-7053.0 / 39385.0, 86551.0 / 94524.0, -46969.0 / 23631.0, 53369.0 / 15754.0, -87904.0 / 23631.0, 820271.0 / 472620.0, -1296.0 / 7877.0, 96.0 / 7877.0, 0.0},
            {21035.0 / 525612.0, -24641.0 / 131403.0, 30409.0 / 87602.0, -54899.0 / 131403.0, 820271.0 / 525612.0, -117600.0 / 43801.0, 64800.0 / 43801.0, -6480.0 / 43801.0, 480.0 / 43801.0}};
        const double bs[5] = {25.0 / 12.0, -4.0, 3.0, -4.0 / 3.0, 1.0 / 4.0};
        fill(6, 9, 5, bh, d, &bd[0][0], bs);
        return true;
    }
    if (p == 8) {
        const double bh[8] = {5080320.0 / 1498139.0, 725760.0 / 1107307.0, 80640.0 / 20761.0, 725760.0 / 1304999.0, 725760.0 / 299527.0, 80640.0 / 103097.0, 725760.0 / 670091.0, 5080320.0 / 5127739.0};
        const double d[9] = {-1.0 / 560.0, 8.0 / 315.0, -1.0 / 5.0, 8.0 / 5.0, -205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
        const double bd[8][12] = {
            {4870382994799.0 / 1358976868290.0, -893640087518.0 / 75498714905.0, 926594825119.0 / 60398971924.0, -1315109406200.0 / 135897686829.0, 39126983272.0 / 15099742981.0, 12344491342.0 / 75498714905.0, -451560522577.0 / 2717953736580.0, 0.0, 0.0, 0.0, 0.0, 0.0},
            {333806012194.0 / 390619153855.0, -154646272029.0 / 111605472530.0, 1168338040.0 / 33481641759.0, 82699112501.0 / 133926567036.0, -171562838.0 / 11160547253.0, -28244698346.0 / 167408208795.0, 11904122576.0 / 167408208795.0, -2598164715.0 / 312495323084.0, 0.0, 0.0, 0.0, 0.0},
            {7838984095.0 / 52731029988.0, 1168338040.0 / 5649753213.0, -88747895.0 / 144865467.0, 423587231.0 / 627750357.0, -43205598281.0 / 22599012852.0, 4876378562.0 / 1883251071.0, -5124426509.0 / 3766502142.0, 10496900965.0 / 39548272491.0, 0.0, 0.0, 0.0, 0.0},
            {-94978241528.0 / 828644350023.0, 82699112501.0 / 157837019052.0, 1270761693.0 / 13153084921.0, -167389605005.0 / 118377764289.0, 48242560214.0 / 39459254763.0, -31673996013.0 / 52612339684.0, 43556319241.0 / 118377764289.0, -44430275135.0 / 552429566682.0, 0.0, 0.0, 0.0, 0.0},
            {1455067816.0 / 21132528431.0, -171562838.0 / 3018932633.0, -43205598281.0 / 36227191596.0, 48242560214.0 / 9056797899.0, -52276055645.0 / 6037865266.0, 57521587238.0 / 9056797899.0, -80321706377.0 / 36227191596.0, 8078087158.0 / 21132528431.0, -1296.0 / 299527.0, 0.0, 0.0, 0.0},
            {10881504334.0 / 327321118845.0, -28244698346.0 / 140280479505.0, 4876378562.0 / 9352031967.0, -10557998671.0 / 12469375956.0, 57521587238.0 / 28056095901.0, -278531401019.0 / 93520319670.0, 73790130002.0 / 46760159835.0, -137529995233.0 / 785570685228.0, 2048.0 / 103097.0, -144.0 / 103097.0, 0.0, 0.0},
            {-135555328849.0 / 8509847458140.0, 11904122576.0 / 101307707835.0, -5124426509.0 / 13507694378.0, 43556319241.0 / 60784624701.0, -80321706377.0 / 81046166268.0, 73790130002.0 / 33769235945.0, -950494905688.0 / 303923123505.0, 239073018673.0 / 141830790969.0, -145152.0 / 670091.0, 18432.0 / 670091.0, -1296.0 / 670091.0, 0.0},
            {0.0, -2598164715.0 / 206729925524.0, 10496900965.0 / 155047444143.0, -44430275135.0 / 310094888286.0, 425162482.0 / 2720130599.0, -137529995233.0 / 620189776572.0, 239073018673.0 / 155047444143.0, -144648000000.0 / 51682481381.0, 8128512.0 / 5127739.0, -1016064.0 / 5127739.0, 129024.0 / 5127739.0, -9072.0 / 5127739.0}};
        const double bs[7] = {4723.0 / 2100.0, -839.0 / 175.0, 157.0 / 35.0, -278.0 / 105.0, 103.0 / 140.0, 1.0 / 175.0, -6.0 / 175.0};
        fill(8, 12, 7, bh, d, &bd[0][0], bs);
        return true;
    }
    if (p == 10) {
        const double bh[11] = {18289152000.0 / 5261271563.0, 1828915200.0 / 2881040311.0, 406425600.0 / 52175551.0, 6096384.0 / 11662993.0, 87091200.0 / 50124587.0, 72576000.0 / 50124587.0, 87091200.0 / 148333439.0, 152409600.0 / 63867949.0, 16257024.0 / 20608675.0, 1828915200.0 / 1704508063.0, 18289152000.0 / 18425967263.0};
        const double d[11] = {1.0 / 3150.0, -5.0 / 1008.0, 5.0 / 126.0, -5.0 / 21.0, 5.0 / 3.0, -5269.0 / 1800.0, 5.0 / 3.0, -5.0 / 21.0, 5.0 / 126.0, -5.0 / 1008.0, 1.0 / 3150.0};
        // upper triangle of M (r <= c, 1-based), diagonal_sbp.jl:319-369
        const struct { int r, c; double v; } mu[66] = {
            {1, 1, 1.2056593789671863908}, {1, 2, -1.3378814169347239658}, {1, 3, 0.0036847309286546532061},
            {1, 4, 0.15698288365600946515}, {1, 5, -0.0037472461482539197952},
            {1, 6, -0.0062491712449361657064}, {1, 7, -0.029164045872729581661},
            {1, 8, 0.00054848184117832929161}, {1, 9, 0.013613461413384884448},
            {1, 10, -0.0025059220258337808220}, {1, 11, -0.00094113457993630916498},
            {2, 2, 2.1749807117105597139}, {2, 3, -0.12369059547124894597}, {2, 4, -0.83712574037924152603},
            {2, 5, 0.050065127254670973258}, {2, 6, 0.0081045853127317536361},
            {2, 7, 0.097405846039248226536}, {2, 8, -0.00068942461520402214720},
            {2, 9, -0.041326971493379188475}, {2, 10, 0.0075778529605774119402},
            {2, 11, 0.0025800256160095691057}, {3, 3, 0.18361596652499065332},
            {3, 4, 0.048289690013342693109}, {3, 5, -0.19719621435164680412}, {3, 6, 0.11406859029505842791},
            {3, 7, -0.029646295985488126964}, {3, 8, -0.0016038463172861201306},
            {3, 9, 0.0032879841528337653050}, {3, 10, -0.00093242311589807387463},
            {3, 11, 0.00012241332668787820533}, {4, 4, 1.2886524606662484673},
            {4, 5, -0.14403037739488789185}, {4, 6, -0.44846291607489015475},
            {4, 7, -0.10598334599408054277}, {4, 8, -0.015873275740355918053},
            {4, 9, 0.073988493386459608166}, {4, 10, -0.012508848749152899785},
            {4, 11, -0.0039290233894513005339}, {5, 5, 0.51482665719685186210},
            {5, 6, 0.051199577887125103015}, {5, 7, -0.36233561810883077365},
            {5, 8, 0.091356850268746392169}, {5, 9, 0.0024195916108052419451},
            {5, 10, -0.0018564214413731389338}, {5, 11, -0.00070192677320704413827},
            {6, 6, 0.68636003380365860083}, {6, 7, -0.28358848290867614908}, {6, 8, -0.13836006478253396528},
            {6, 9, 0.0076158070663111995297}, {6, 10, 0.011447010307180005164},
            {6, 11, -0.0021349696610286552676}, {7, 7, 1.5216081480839085990},
            {7, 8, -0.42653865162216293237}, {7, 9, -0.42047484981879143123},
            {7, 10, 0.019813359263872926304}, {7, 11, 0.019221397241190103344},
            {8, 8, 1.0656733504627815335}, {8, 9, -0.66921872668484232217}, {8, 10, 0.12022033144141336599},
            {8, 11, -0.030157881394591483631}, {9, 9, 2.4064247712949611684},
            {9, 10, -1.5150200315922263367}, {9, 11, 0.17373015320416595052},
            {10, 10, 2.7682502485427255096}, {10, 11, -1.5975407111468405444},
            {11, 11, 2.9033627686681129471},
        };
        const double bsneg[7] = {-49.0 / 20.0, 6.0, -15.0 / 2.0, 20.0 / 3.0, -15.0 / 4.0, 6.0 / 5.0, -1.0 / 6.0};
        // M is 11 x 16: the symmetric 11 x 11 block from its upper triangle, and
        // rows 7..11 reaching columns 12..16 with the negated interior stencil;
        // then bd = diag(bhinv) (-M + e0 BS) with BS = -bsneg.
        double M[11][16] = {};
        for (const auto& e : mu) {
            M[e.r - 1][e.c - 1] = e.v;
            M[e.c - 1][e.r - 1] = e.v;
        }
        for (int k = 1; k <= 5; ++k)
            for (int q = 1; q <= k; ++q) M[5 + k][10 + q] = -d[k - q];
        double bs[7], bd[11][16];
        for (int c = 0; c < 7; ++c) bs[c] = -bsneg[c];
        for (int r = 0; r < 11; ++r)
            for (int c = 0; c < 16; ++c) bd[r][c] = bh[r] * (-M[r][c] + (r == 0 && c < 7 ? bs[c] : 0.0));
        fill(11, 16, 7, bh, d, &bd[0][0], bs);
        return true;
    }
    return false;
}

// Dense dimensionless factors (hw, m, k) for side n (n >= 2*bm, n >= bn).
inline void dense_factors(const Sbp1D& s, int n, std::vector<double>& hw, std::vector<double>& m,
                          std::vector<double>& k) {
    std::vector<double> D2(n * n, 0.0), BS(n * n, 0.0);
    const int half = s.p / 2;
    for (int r = s.bm; r < n - s.bm; ++r)
        for (int t = 0; t <= s.p; ++t) D2[r * n + (r - half + t)] = s.d[t];
    for (int r = 0; r < s.bm; ++r)
        for (int c = 0; c < s.bn; ++c) {
            D2[r * n + c] = s.bd[r][c];
            D2[(n - 1 - r) * n + (n - 1 - c)] = s.bd[r][c];
        }
    for (int c = 0; c < s.nbs; ++c) {
        BS[0 * n + c] = s.bs[c];                  // -S0 row 1
        BS[(n - 1) * n + (n - 1 - c)] = s.bs[c];  //  SN row N
    }
    hw.assign(n, 1.0);
    for (int r = 0; r < s.bm; ++r) {
        hw[r] = 1.0 / s.bhinv[r];
        hw[n - 1 - r] = 1.0 / s.bhinv[r];
    }
    m.assign(n * n, 0.0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) m[r * n + c] = BS[r * n + c] - hw[r] * D2[r * n + c];
    k = m;
    for (int c = 0; c < n; ++c) {
        k[0 * n + c] -= BS[0 * n + c];
        k[(n - 1) * n + c] -= BS[(n - 1) * n + c];
        k[c * n + 0] -= BS[0 * n + c];
        k[c * n + (n - 1)] -= BS[(n - 1) * n + c];
    }
    k[0] += 13.0;
    k[(n - 1) * n + (n - 1)] += 13.0;
}

// Fills the kernel table for side N. Returns false when N is too small.
inline bool build_optab(int p, int N, OpTab* t, RhsTab* rt) {
    Sbp1D s;
    if (!sbp_1d(p, &s)) return false;
    // closure rows / columns of the factors (rows below R are the interior stencil)
    // R = max(bm, nbs) (the SAT column reaches nbs rows); C = max(bn, R + p/2)
    const int R = s.bm > s.nbs ? s.bm : s.nbs;
    const int C = s.bn > R + p / 2 ? s.bn : R + p / 2;
    if (R > kMaxR || C > kMaxC) return false;
    if (N < 2 * C || N < 2 * s.bm || N < s.bn) return false;
    const int n0 = 2 * C + 4;
    std::vector<double> hw, m, k;
    dense_factors(s, n0, hw, m, k);
    std::memset(t, 0, sizeof(*t));
    t->N = N; t->R = R; t->C = C; t->RH = s.bm; t->H = p / 2;
    for (int q = 0; q <= p / 2; ++q) t->cint[q] = -s.d[p / 2 + q];
    // every row from R on is the interior stencil (checked on the dense build)
    for (int r = R; r < n0 - R; ++r)
        for (int c = 0; c < n0; ++c) {
            const int o = c - r < 0 ? r - c : c - r;
            const double want = o <= p / 2 ? t->cint[o] : 0.0;
            if (std::fabs(m[r * n0 + c] - want) > 1e-12 || std::fabs(k[r * n0 + c] - want) > 1e-12) return false;
        }
    for (int r = 0; r < s.bm; ++r) t->hw[r] = hw[r];
    for (int r = 0; r < R; ++r) {
        for (int c = 0; c < C; ++c) {
            t->mb[r][c] = m[r * n0 + c];
            t->kb[r][c] = k[r * n0 + c];
        }
        for (int c = C; c < n0; ++c)
            if (m[r * n0 + c] != 0.0 || k[r * n0 + c] != 0.0) return false;
        t->md[r] = m[r * n0 + r];
        t->kd[r] = k[r * n0 + r];
        double ms = 0.0, ks = 0.0, mmax = 0.0;
        for (int c = 0; c < C; ++c) {
            ms += t->mb[r][c];
            ks += t->kb[r][c];
            mmax = std::fmax(mmax, std::fabs(t->mb[r][c]));
        }
        if (std::fabs(ms) > 1e-12 * mmax) return false;   // m 1 = 0 must hold
        t->mrs[r] = 0.0;
        t->krs[r] = ks;
    }
    t->mid = -s.d[p / 2];
    t->idint = 1.0 / (2.0 * t->mid);
    if (rt) {
        std::memset(rt, 0, sizeof(*rt));
        rt->nbs = s.nbs;
        for (int c = 0; c < s.nbs; ++c) rt->bs[c] = s.bs[c];
        rt->h = 1.0 / (N - 1);
        rt->bW = rt->bE = 1.0;
    }
    return true;
}

// Dense assembled A (row-major, n = N*N) for a small side; used for the
// coarsest-level factorisation.
inline void dense_operator(const OpTab& t, std::vector<double>& A) {
    const int N = t.N, n = N * N;
    A.assign((size_t)n * n, 0.0);
    auto hwf = [&](int r) {
        if (r < t.RH) return t.hw[r];
        if (r >= N - t.RH) return t.hw[N - 1 - r];
        return 1.0;
    };
    const int H = t.H;
    auto coef = [&](bool is_k, int r, int c) -> double {
        if (r < t.R) return (c < t.C) ? (is_k ? t.kb[r][c] : t.mb[r][c]) : 0.0;
        if (r >= N - t.R) {
            const int rr = N - 1 - r, cc = N - 1 - c;
            return (cc >= 0 && cc < t.C) ? (is_k ? t.kb[rr][cc] : t.mb[rr][cc]) : 0.0;
        }
        const int s = c - r < 0 ? r - c : c - r;
        return s <= H ? t.cint[s] : 0.0;
    };
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const int row = i + j * N;
            for (int c = 0; c < N; ++c) {
                const double kv = coef(true, j, c);
                if (kv != 0.0) A[(size_t)row * n + (i + c * N)] += hwf(i) * kv;
                const double mv = coef(false, i, c);
                if (mv != 0.0) A[(size_t)row * n + (c + j * N)] += hwf(j) * mv;
            }
        }
}

// In-place inverse of an SPD matrix via Cholesky (row-major n x n).
inline bool spd_inverse(std::vector<double>& A, int n) {
    std::vector<double> L((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double v = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = v / ljj;
        }
    }
    // Solve L L^T X = I for all columns at once, a row at a time (the inner
    // loops run along contiguous rows): Y = L^-1 (lower triangular, so row i
    // of Y has entries in columns <= i only), then X = L^-T Y.
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double* yi = &X[(size_t)i * n];
        yi[i] = 1.0;
        for (int k = 0; k < i; ++k) {
            const double l = L[(size_t)i * n + k];
            const double* yk = &X[(size_t)k * n];
            for (int c = 0; c <= k; ++c) yi[c] -= l * yk[c];
        }
        const double inv = 1.0 / L[(size_t)i * n + i];
        for (int c = 0; c <= i; ++c) yi[c] *= inv;
    }
    for (int i = n - 1; i >= 0; --i) {
        double* xi = &X[(size_t)i * n];
        for (int k = i + 1; k < n; ++k) {
            const double l = L[(size_t)k * n + i];
            const double* xk = &X[(size_t)k * n];
            for (int c = 0; c < n; ++c) xi[c] -= l * xk[c];
        }
        const double inv = 1.0 / L[(size_t)i * n + i];
        for (int c = 0; c < n; ++c) xi[c] *= inv;
    }
    // Symmetrise (the inverse of a symmetric matrix).
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double v = 0.5 * (X[(size_t)i * n + j] + X[(size_t)j * n + i]);
            X[(size_t)i * n + j] = X[(size_t)j * n + i] = v;
        }
    A.swap(X);
    return true;
}

}  // namespace sbp

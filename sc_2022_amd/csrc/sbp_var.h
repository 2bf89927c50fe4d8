// sbp_var.h -- the variable-coefficient operator, orders 2, 4 and 6.
//
// d/dx (bx d/dx) + d/dy (by d/dy) with one coefficient profile per axis
// (bx along the slow, Dirichlet index j; by along the fast, Neumann index i),
// from the reference's variable_diagonal_sbp_D2 (diagonal_sbp.jl:470-740)
// used on both axes as its get_operators does (legacy_functions.jl:100-174).
// The tensor form of the constant case carries over:
//     (A u)(i,j) = hw[i] sum_j' k[j][j'] u(i,j') + hw[j] sum_i' m[i][i'] u(i',j)
// with m = h M(by) (the reference's M before its division by h) and
// k = m(bx) - (E bs + bs^T E) + 13 b_face E, bs the b_face-weighted boundary
// row and the Dirichlet penalty scaled by the face coefficient (the
// reference's tau = -13/h at b = 1).  The factors are no longer Toeplitz:
// every interior row has its own 2H+1 coefficients (stored offset-major,
// N per offset, so a wavefront reads them coalesced), and the two faces have
// different closure blocks.  H weights do not depend on b.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

#include "sbp_kernels.h"
#include "sbp_var_tab.h"

namespace sbp {

// ---------------------------------------------------------------------------
// Host: the entries of M(b) (dimensionless, i.e. h M), restated from
// diagonal_sbp.jl:470-566 with 1-based indices as the reference writes them.
// ---------------------------------------------------------------------------
// p = 4 boundary block: entry (r, c), r <= c, is sum num/den * b_k; the
// integer literals are divided in double precision as the reference does.
// The p = 6 block's weights are decimals (den = 1).
struct VarTerm { double num, den; int k; };
struct VarEntry { int r, c; VarTerm t[9]; int nt; };

inline const std::vector<VarEntry>& var_p4_block() {
    static const std::vector<VarEntry> e = {
        {1, 1, {{12, 17, 1}, {59, 192, 2}, {27010400129.0, 345067064608.0, 3}, {69462376031.0, 2070402387648.0, 4}}, 4},
        {1, 2, {{-59, 68, 1}, {-6025413881.0, 21126554976.0, 3}, {-537416663.0, 7042184992.0, 4}}, 3},
        {1, 3, {{2, 17, 1}, {-59, 192, 2}, {213318005.0, 16049630912.0, 4}, {2083938599.0, 8024815456.0, 3}}, 4},
        {1, 4, {{3, 68, 1}, {-1244724001.0, 21126554976.0, 3}, {752806667.0, 21126554976.0, 4}}, 3},
        {1, 5, {{49579087.0, 10149031312.0, 3}, {-49579087.0, 10149031312.0, 4}}, 2},
        {1, 6, {{-1, 784, 4}, {1, 784, 3}}, 2},
        {2, 2, {{3481, 3264, 1}, {9258282831623875.0, 7669235228057664.0, 3}, {236024329996203.0, 1278205871342944.0, 4}}, 3},
        {2, 3, {{-59, 408, 1}, {-29294615794607.0, 29725717938208.0, 3}, {-2944673881023.0, 29725717938208.0, 4}}, 3},
        {2, 4, {{-59, 1088, 1}, {260297319232891.0, 2556411742685888.0, 3}, {-60834186813841.0, 1278205871342944.0, 4}}, 3},
        {2, 5, {{-1328188692663.0, 37594290333616.0, 3}, {1328188692663.0, 37594290333616.0, 4}}, 2},
        {2, 6, {{-8673, 2904112, 3}, {8673, 2904112, 4}}, 2},
        {3, 3, {{1, 51, 1}, {59, 192, 2}, {13777050223300597.0, 26218083221499456.0, 4}, {564461, 13384296, 5},
                {378288882302546512209.0, 270764341349677687456.0, 3}}, 5},
        {3, 4, {{1, 136, 1}, {-125059, 743572, 5}, {-4836340090442187227.0, 5525802884687299744.0, 3},
                {-17220493277981.0, 89177153814624.0, 4}}, 4},
        {3, 5, {{-10532412077335.0, 42840005263888.0, 4}, {1613976761032884305.0, 7963657098519931984.0, 3},
                {564461, 4461432, 5}}, 3},
        {3, 6, {{-960119, 1280713392, 4}, {-3391, 6692148, 5}, {33235054191.0, 26452850508784.0, 3}}, 3},
        {4, 4, {{3, 1088, 1}, {507284006600757858213.0, 475219048083107777984.0, 3}, {1869103, 2230716, 5}, {1, 24, 6},
                {1950062198436997.0, 3834617614028832.0, 4}}, 5},
        {4, 5, {{-4959271814984644613.0, 20965546238960637264.0, 3}, {-1, 6, 6}, {-15998714909649.0, 37594290333616.0, 4},
                {-375177, 743572, 5}}, 4},
        {4, 6, {{-368395, 2230716, 5}, {752806667.0, 539854092016.0, 3}, {1063649, 8712336, 4}, {1, 8, 6}}, 4},
        {5, 5, {{8386761355510099813.0, 128413970713633903242.0, 3}, {2224717261773437.0, 2763180339520776.0, 4},
                {5, 6, 6}, {1, 24, 7}, {280535, 371786, 5}}, 5},
        {5, 6, {{-35039615, 213452232, 4}, {-1, 6, 7}, {-13091810925.0, 13226425254392.0, 3}, {-1118749, 2230716, 5},
                {-1, 2, 6}}, 5},
        {6, 6, {{3290636, 80044587, 4}, {5580181, 6692148, 5}, {5, 6, 7}, {1, 24, 8},
                {660204843.0, 13226425254392.0, 3}, {3, 4, 6}}, 6},
    };
    return e;
}

// p = 6 boundary block (diagonal_sbp.jl:577-624), b1..b12.
inline const std::vector<VarEntry>& var_p6_block() {
    static const std::vector<VarEntry> e = {
        {1, 1, {{0.79126675946955820939, 1, 1}, {0.29684720906380007429, 1, 2}, {0.0031855190887964290152, 1, 3}, {0.016324040425909519534, 1, 4}, {0.031603022440944150877, 1, 5}, {0.031679647480161052996, 1, 6}, {0.031485777339472539205, 1, 7}}, 7},
        {1, 2, {{-1.0166893393503381444, 1, 1}, {-0.028456273704916113690, 1, 3}, {-0.041280298383492988198, 1, 4}, {-0.13922814516201405075, 1, 5}, {-0.11957773256112017666, 1, 6}, {-0.11942677565293334109, 1, 7}}, 6},
        {1, 3, {{0.070756429372437150463, 1, 1}, {-0.18454761060241510503, 1, 2}, {-0.043641631471118923470, 1, 4}, {0.24323679072077324609, 1, 5}, {0.15821270735372154440, 1, 6}, {0.16023485783647863076, 1, 7}}, 6},
        {1, 4, {{0.22519915328913532127, 1, 1}, {-0.16627487110970548953, 1, 2}, {0.027105309616486712977, 1, 3}, {-0.19166461859684399091, 1, 5}, {-0.076841171601990145944, 1, 6}, {-0.082195869498316975759, 1, 7}}, 6},
        {1, 5, {{-0.052244034642020563167, 1, 1}, {0.044400639485098762210, 1, 2}, {-0.0010239765473093878745, 1, 3}, {0.074034846453161740905, 1, 4}, {0.012416255689984968954, 1, 6}, {0.071886528478926012827, 1, 5}, {0.013793629971047355034, 1, 7}}, 7},
        {1, 6, {{-0.018288968138771973527, 1, 1}, {0.0095746331632217580607, 1, 2}, {-0.00081057845305764042779, 1, 3}, {-0.0073488455877755196984, 1, 4}, {0.010636019497239069970, 1, 5}, {-0.013159670383826183824, 1, 6}, {-0.021179364788387535246, 1, 7}}, 7},
        {1, 7, {{0.0019118885633161709274, 1, 4}, {-0.040681303555291499361, 1, 5}, {0.013196749810737491670, 1, 6}, {0.025572665181237836763, 1, 7}}, 4},
        {1, 8, {{0.015596528711367857640, 1, 5}, {-0.0064861841573315378995, 1, 6}, {-0.0091103445540363197401, 1, 7}}, 3},
        {1, 9, {{0.00055939836966298630593, 1, 6}, {-0.0013848225351007963723, 1, 5}, {0.00082542416543781006633, 1, 7}}, 3},
        {2, 2, {{1.3063321571116676286, 1, 1}, {0.25420017604573457435, 1, 3}, {0.10438978280925626095, 1, 4}, {0.66723280210321129509, 1, 5}, {0.46818193597227494411, 1, 6}, {0.46764154101958369201, 1, 7}}, 6},
        {2, 3, {{-0.090914102699924646049, 1, 1}, {0.11036113131714764253, 1, 4}, {-1.2903975449975188870, 1, 5}, {-0.66396052487350447871, 1, 6}, {-0.66159744640052061842, 1, 7}}, 5},
        {2, 4, {{-0.28935573956534316666, 1, 1}, {-0.24213200040645927216, 1, 3}, {1.1876702550280310277, 1, 5}, {0.39565981499041363328, 1, 6}, {0.38600489217558000007, 1, 7}}, 5},
        {2, 5, {{0.067127744758037639890, 1, 1}, {0.0091471926820756301800, 1, 3}, {-0.18721961430038080217, 1, 4}, {-0.13193585588531745301, 1, 6}, {-0.48715757368119118874, 1, 5}, {-0.10475163122754481381, 1, 7}}, 6},
        {2, 6, {{0.023499279745900688694, 1, 1}, {0.0072409053835651813164, 1, 3}, {0.018583789963916794487, 1, 4}, {-0.092896161339386761743, 1, 5}, {0.12235132704188076670, 1, 6}, {0.11135203204362950339, 1, 7}}, 6},
        {2, 7, {{-0.0048347914064469075906, 1, 4}, {0.23106838326878204031, 1, 5}, {-0.10807741421960079917, 1, 6}, {-0.11815617764273433354, 1, 7}}, 4},
        {2, 8, {{-0.083681414344034553537, 1, 5}, {0.040934994667670546616, 1, 6}, {0.042746419676364006921, 1, 7}}, 3},
        {2, 9, {{-0.0035765451326969831434, 1, 6}, {0.0073893991241210786821, 1, 5}, {-0.0038128539914240955387, 1, 7}}, 3},
        {3, 3, {{0.0063271611471368738078, 1, 1}, {0.11473182007158685275, 1, 2}, {0.11667405542796800075, 1, 4}, {2.7666108082854440372, 1, 5}, {1.0709206899608171042, 1, 6}, {1.0131613910329730572, 1, 7}}, 6},
        {3, 4, {{0.020137694138847972466, 1, 1}, {0.10337179946308864017, 1, 2}, {-2.9132216211517427243, 1, 5}, {-0.87558073434822622598, 1, 6}, {-0.69099571834888124265, 1, 7}}, 5},
        {3, 5, {{-0.0046717510915754628683, 1, 1}, {-0.027603533656377128278, 1, 2}, {-0.19792902986208699745, 1, 4}, {0.54029853383734330523, 1, 6}, {1.2391775930319110779, 1, 5}, {0.26280380502473582273, 1, 7}}, 6},
        {3, 6, {{-0.0016354308669218878195, 1, 1}, {-0.0059524752758832596197, 1, 2}, {0.019646827777442752194, 1, 4}, {0.32366400126390466006, 1, 5}, {-0.46595166932288709739, 1, 6}, {-0.22172727209417368594, 1, 7}}, 6},
        {3, 7, {{-0.0051113531893524745496, 1, 4}, {-0.53558781637747543460, 1, 5}, {0.33283351044897389336, 1, 6}, {0.20786565911785401579, 1, 7}}, 4},
        {3, 8, {{0.18243281741342895622, 1, 5}, {-0.10598160301968184459, 1, 6}, {-0.076451214393747111630, 1, 7}}, 3},
        {3, 9, {{0.0092090899634437994856, 1, 6}, {-0.015915028188724931671, 1, 5}, {0.0067059382252811321853, 1, 7}}, 3},
        {4, 4, {{0.064092997759871869867, 1, 1}, {0.093136576388046999489, 1, 2}, {0.23063676246347492291, 1, 3}, {3.6894403082837166203, 1, 5}, {1.1905503386876088738, 1, 6}, {0.59124795468888565194, 1, 7}}, 6},
        {4, 5, {{-0.014868958192656041286, 1, 1}, {-0.024870405993901607642, 1, 2}, {-0.0087129289077117541871, 1, 3}, {-1.2635078373718242057, 1, 6}, {-0.30583173978439973269, 1, 7}, {-1.4706919260458029548, 1, 5}}, 6},
        {4, 6, {{-0.0052051474298559556576, 1, 1}, {-0.0053630987475285424890, 1, 2}, {-0.0068971427657906095463, 1, 3}, {-0.78575245216674501017, 1, 5}, {0.22911480054237346001, 1, 7}, {0.99770643562927505292, 1, 6}}, 6},
        {4, 7, {{0.66972974880676622652, 1, 5}, {-0.50132473560721279390, 1, 6}, {-0.17951612431066454373, 1, 7}}, 3},
        {4, 8, {{-0.20229090601117515652, 1, 5}, {0.14534218580636584986, 1, 6}, {0.056948720204809306656, 1, 7}}, 3},
        {4, 9, {{-0.012004296184410038337, 1, 6}, {-0.0047769156693859238415, 1, 7}, {0.016781211853795962179, 1, 5}}, 3},
        {5, 5, {{0.0034494550959102336252, 1, 1}, {0.0066411834994278261016, 1, 2}, {0.00032915450832718628585, 1, 3}, {0.33577217075764772000, 1, 4}, {2.0964133295790264390, 1, 6}, {0.23173232041831268550, 1, 7}, {0.0061078257643682645765, 1, 8}, {0.71091258506833766956, 1, 5}}, 8},
        {5, 6, {{0.0012075440723041938061, 1, 1}, {0.0014321166657521476075, 1, 2}, {0.00026055826461832559573, 1, 3}, {-0.033329411132516353908, 1, 4}, {-0.28082416973855326836, 1, 7}, {-0.027209080835250836084, 1, 8}, {0.10458654356829219874, 1, 5}, {-1.3484369866671155432, 1, 6}}, 8},
        {5, 7, {{0.0086710380841746926251, 1, 4}, {0.17360734113554285637, 1, 6}, {0.053313621252876254126, 1, 8}, {-0.24249352624045263018, 1, 5}, {0.15690152576785882706, 1, 7}}, 5},
        {5, 8, {{-0.086316839802171222760, 1, 6}, {0.026988423604709992435, 1, 7}, {0.080981941477156510853, 1, 5}, {-0.032764636390806391639, 1, 8}}, 4},
        {5, 9, {{0.0074620594845308550733, 1, 6}, {-0.00081216403616686789496, 1, 7}, {0.00055227020881270902093, 1, 8}, {-0.0072021656571766961993, 1, 5}}, 4},
        {6, 6, {{0.00042272261734493450425, 1, 1}, {0.00030882419443789644048, 1, 2}, {0.00020625757066474306202, 1, 3}, {0.0033083434042009682567, 1, 4}, {0.58280470164050018158, 1, 5}, {0.80541742203662154736, 1, 7}, {0.13383632334100334433, 1, 8}, {0.0055555555555555555556, 1, 9}, {1.1903620718618930511, 1, 6}}, 9},
        {6, 7, {{-0.00086070442526864133026, 1, 4}, {-0.17480747086739049893, 1, 5}, {-0.31325448501150501650, 1, 8}, {-0.025000000000000000000, 1, 9}, {-0.31691663053104292713, 1, 7}, {-0.66916070916479291611, 1, 6}}, 6},
        {6, 8, {{0.033546617916933521087, 1, 5}, {-0.33436200223869714050, 1, 7}, {0.050000000000000000000, 1, 9}, {0.21697906098076027508, 1, 6}, {0.18383632334100334433, 1, 8}}, 5},
        {6, 9, {{0.029125184768230046430, 1, 7}, {0.022790919164749163916, 1, 8}, {-0.030689859975187405305, 1, 6}, {-0.0017817995133473605962, 1, 5}, {-0.030555555555555555556, 1, 9}}, 5},
        {7, 7, {{0.00022392237357715991790, 1, 4}, {0.12754377854309566738, 1, 5}, {1.0116994839296081646, 1, 6}, {0.96988172751725752475, 1, 8}, {0.12500000000000000000, 1, 9}, {0.0055555555555555555556, 1, 10}, {0.48231775430312815001, 1, 7}}, 7},
        {7, 8, {{-0.037841139730330129499, 1, 5}, {-0.29975568851348273616, 1, 6}, {-0.30000000000000000000, 1, 9}, {-0.025000000000000000000, 1, 10}, {-0.39914868674468211784, 1, 7}, {-0.43825448501150501650, 1, 8}}, 6},
        {7, 9, {{0.046981462180226839339, 1, 6}, {-0.29668637874712374587, 1, 8}, {0.050000000000000000000, 1, 10}, {0.17163557041460064817, 1, 7}, {0.0030693461522962583624, 1, 5}, {0.17500000000000000000, 1, 9}}, 6},
        {8, 8, {{0.012303289427168044554, 1, 5}, {0.11836475296458983325, 1, 6}, {0.94105118982279433342, 1, 7}, {0.95000000000000000000, 1, 9}, {0.12500000000000000000, 1, 10}, {0.0055555555555555555556, 1, 11}, {0.56994743445211445545, 1, 8}}, 7},
        {8, 9, {{-0.023080678926719163396, 1, 6}, {-0.29866250537751494972, 1, 7}, {-0.30000000000000000000, 1, 10}, {-0.025000000000000000000, 1, 11}, {-0.0010477348605150508026, 1, 5}, {-0.42720908083525083608, 1, 8}, {-0.42500000000000000000, 1, 9}}, 7},
        {9, 9, {{0.0051393702211491099770, 1, 6}, {0.12477232150094220014, 1, 7}, {0.95055227020881270902, 1, 8}, {0.95000000000000000000, 1, 10}, {0.12500000000000000000, 1, 11}, {0.0055555555555555555556, 1, 12}, {0.000091593624651536418269, 1, 5}, {0.56111111111111111111, 1, 9}}, 8},
    };
    return e;
}

// Calls put(row, col, value) (1-based) for every stored entry of h M(b)
// (duplicates add, as in the reference's sparse()).
template <class PUT>
inline void var_m_entries(int p, int Np, const double* B, PUT&& put) {
    const int N = Np - 1;
    auto b = [B](int k) { return B[k - 1]; };
    if (p == 2) {
        put(1, 1, (b(1) + b(2)) / 2);
        put(N + 1, N + 1, (b(N) + b(N + 1)) / 2);
        for (int i = 1; i <= N; ++i) put(i + 1, i, -(b(i) + b(i + 1)) / 2);
        for (int i = 2; i <= N; ++i) put(i, i, (b(i - 1) + 2 * b(i) + b(i + 1)) / 2);
        for (int i = 1; i <= N; ++i) put(i, i + 1, -(b(i) + b(i + 1)) / 2);
        return;
    }
    const int nb = p == 4 ? 8 : 12;
    double blo[12], bhi[12];
    for (int k = 0; k < nb; ++k) { blo[k] = B[k]; bhi[k] = B[N - k]; }
    for (const VarEntry& e : (p == 4 ? var_p4_block() : var_p6_block())) {
        double vl = 0.0, vh = 0.0;
        for (int q = 0; q < e.nt; ++q) {
            const double w = e.t[q].num / e.t[q].den;
            vl += w * blo[e.t[q].k - 1];
            vh += w * bhi[e.t[q].k - 1];
        }
        put(e.r, e.c, vl);
        if (e.r != e.c) put(e.c, e.r, vl);
        put(N + 2 - e.r, N + 2 - e.c, vh);
        if (e.r != e.c) put(N + 2 - e.c, N + 2 - e.r, vh);
    }
    if (p == 6) {   // bands by row I (diagonal_sbp.jl:711-728)
        for (int I = 10; I <= N - 5; ++I)
            put(I, I - 3, -(11.0 / 360) * b(I - 3) + (1.0 / 40) * b(I - 2) + (1.0 / 40) * b(I - 1) - (11.0 / 360) * b(I));
        for (int I = 10; I <= N - 6; ++I)
            put(I, I - 2, (1.0 / 20) * b(I - 3) + (7.0 / 40) * b(I - 2) - (3.0 / 10) * b(I - 1) + (7.0 / 40) * b(I) +
                              (1.0 / 20) * b(I + 1));
        for (int I = 10; I <= N - 7; ++I)
            put(I, I - 1, -(1.0 / 40) * b(I - 3) - (3.0 / 10) * b(I - 2) - (17.0 / 40) * b(I - 1) -
                              (17.0 / 40) * b(I) - (3.0 / 10) * b(I + 1) - (1.0 / 40) * b(I + 2));
        for (int I = 10; I <= N - 8; ++I)
            put(I, I, (1.0 / 180) * b(I - 3) + (1.0 / 8) * b(I - 2) + (19.0 / 20) * b(I - 1) + (101.0 / 180) * b(I) +
                          (19.0 / 20) * b(I + 1) + (1.0 / 8) * b(I + 2) + (1.0 / 180) * b(I + 3));
        for (int I = 9; I <= N - 8; ++I)
            put(I, I + 1, -(1.0 / 40) * b(I - 2) - (3.0 / 10) * b(I - 1) - (17.0 / 40) * b(I) -
                              (17.0 / 40) * b(I + 1) - (3.0 / 10) * b(I + 2) - (1.0 / 40) * b(I + 3));
        for (int I = 8; I <= N - 8; ++I)
            put(I, I + 2, (1.0 / 20) * b(I - 1) + (7.0 / 40) * b(I) - (3.0 / 10) * b(I + 1) + (7.0 / 40) * b(I + 2) +
                              (1.0 / 20) * b(I + 3));
        for (int I = 7; I <= N - 8; ++I)
            put(I, I + 3, -(11.0 / 360) * b(I) + (1.0 / 40) * b(I + 1) + (1.0 / 40) * b(I + 2) - (11.0 / 360) * b(I + 3));
        return;
    }
    for (int J = 5; J <= N - 5; ++J) put(J + 2, J, (1.0 / 8) * b(J + 2) - (1.0 / 6) * b(J + 1) + (1.0 / 8) * b(J));
    for (int J = 6; J <= N - 5; ++J)
        put(J + 1, J, -(1.0 / 6) * b(J + 2) - (1.0 / 2) * b(J + 1) - (1.0 / 2) * b(J) - (1.0 / 6) * b(J - 1));
    for (int J = 7; J <= N - 5; ++J)
        put(J, J, (1.0 / 24) * b(J + 2) + (5.0 / 6) * b(J + 1) + (3.0 / 4) * b(J) + (5.0 / 6) * b(J - 1) +
                      (1.0 / 24) * b(J - 2));
    for (int J = 7; J <= N - 4; ++J)
        put(J - 1, J, -(1.0 / 6) * b(J + 1) - (1.0 / 2) * b(J) - (1.0 / 2) * b(J - 1) - (1.0 / 6) * b(J - 2));
    for (int J = 7; J <= N - 3; ++J) put(J - 2, J, (1.0 / 8) * b(J) - (1.0 / 6) * b(J - 1) + (1.0 / 8) * b(J - 2));
}

// Builds one axis' factor (m, or k when dirichlet) from the coefficient
// profile b (N values): closure blocks into ax, the band (2H+1) x N into
// band.  R / C: closure rows / columns (var_R, var_C).
// Returns false when N is too small or an entry falls outside the layout.
inline bool build_var_axis(int p, int N, const double* b, bool dirichlet, VarAxis* ax, std::vector<double>& band) {
    const int H = p / 2, R = var_R(p), C = var_C(p);
    if (N < 2 * C) return false;
    std::memset(ax->lo, 0, sizeof(ax->lo));
    std::memset(ax->hi, 0, sizeof(ax->hi));
    band.assign((size_t)(2 * H + 1) * N, 0.0);
    bool ok = true;
    auto put0 = [&](int r, int c, double v) {   // 0-based
        if (r < R) {
            if (c >= C) { ok = false; return; }
            ax->lo[r][c] += v;
        } else if (r >= N - R) {
            const int rr = N - 1 - r, cc = N - 1 - c;
            if (cc < 0 || cc >= C) { ok = false; return; }
            ax->hi[rr][cc] += v;
        } else {
            const int s = c - r;
            if (s < -H || s > H) { ok = false; return; }
            band[(size_t)(s + H) * N + r] += v;
        }
    };
    var_m_entries(p, N, b, [&](int r, int c, double v) { put0(r - 1, c - 1, v); });
    if (dirichlet) {
        // k = m - (E bs + bs^T E) + 13 b_face E, bs = b_face * BS (dimensionless)
        const double bs2[3] = {3.0 / 2.0, -2.0, 1.0 / 2.0};
        const double bs4[4] = {11.0 / 6.0, -3.0, 3.0 / 2.0, -1.0 / 3.0};
        const double bs6[5] = {25.0 / 12.0, -4.0, 3.0, -4.0 / 3.0, 1.0 / 4.0};
        const double* bs = p == 2 ? bs2 : (p == 4 ? bs4 : bs6);
        const int nbs = p == 2 ? 3 : (p == 4 ? 4 : 5);
        for (int c = 0; c < nbs; ++c) {
            put0(0, c, -b[0] * bs[c]);
            put0(c, 0, -b[0] * bs[c]);
            put0(N - 1, N - 1 - c, -b[N - 1] * bs[c]);
            put0(N - 1 - c, N - 1, -b[N - 1] * bs[c]);
        }
        put0(0, 0, 13.0 * b[0]);
        put0(N - 1, N - 1, 13.0 * b[N - 1]);
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------
// sum_c a[r][c] u(c) along one axis for row r (load(c) gives u at index c;
// band: the axis' coefficient band, passed where the kernel has it as an
// argument so that its loads are known to address global memory).
template <int H, class LOAD>
__device__ __forceinline__ double var_row(const VarAxis& a, const double* band, int R, int C, int N, int r,
                                          LOAD&& load) {
    double s = 0.0;
    if (r < R) {
        for (int c = 0; c < C; ++c) s += a.lo[r][c] * load(c);
    } else if (r >= N - R) {
        const int rr = N - 1 - r;
        for (int c = 0; c < C; ++c) s += a.hi[rr][c] * load(N - 1 - c);
    } else {
#pragma unroll
        for (int q = -H; q <= H; ++q) s += band[(size_t)(q + H) * N + r] * load(r + q);
    }
    return s;
}

// A block: kVarThreads threads along i, kVarRows consecutive rows j.
constexpr int kVarThreads = 256;
constexpr int kVarRows = 8;

// Diagonal entry of one axis' factor at row r.
__device__ __forceinline__ double var_diag1(const VarAxis& a, const double* band, int R, int H, int N, int r) {
    if (r < R) return a.lo[r][r];
    if (r >= N - R) return a.hi[N - 1 - r][N - 1 - r];
    return band[(size_t)H * N + r];
}

__device__ __forceinline__ double var_hw(const VarTab& t, int N, int r) {
    const int rr = r < N - 1 - r ? r : N - 1 - r;
    return rr < t.RH ? t.hw[rr] : 1.0;
}

// diag(A) at (i, j): hw_i k_jj + hw_j m_ii.
__device__ __forceinline__ double var_diag(const VarTab& t, const double* kb, const double* mb, int N, int i, int j) {
    return var_hw(t, N, i) * var_diag1(t.k, kb, t.R, t.H, N, j) + var_hw(t, N, j) * var_diag1(t.m, mb, t.R, t.H, N, i);
}

// ---------------------------------------------------------------------------
// Operand (IN) and output (OUT) policies of the variable-coefficient pass.
// The arithmetic is FP64 whatever the element type T of the grid functions.
// IN: the operand at (i, j) as a raw fetch (loads only) and a transform
// make(raw, i, j, d), d = diag(A)(i, j) when kDiag (else unused); kCol: a
// column loader for the interior tiles.  OUT: pre(q) loads what emit needs
// at point q (issued before the block's barrier); emit(q, u, Au, dinv, aux,
// acc) writes; kDiag: emit needs dinv = 1 / diag(A)(i, j); kDot: emit
// accumulates the block partial.  (make likewise receives dinv: a pass
// forms the reciprocal once per point and both sweeps of the first
// smoothing pass multiply by it.)
// ---------------------------------------------------------------------------
template <typename T> struct VarIn {            // u
    const T* u;
    static constexpr bool kDiag = false, kCol = false;
    using Raw = double;
    __device__ __forceinline__ Raw fetch(int i, int j, int N) const { return (double)u[gidx(i, j, N)]; }
    __device__ __forceinline__ double make(const Raw& r, int, int, double) const { return r; }
};

// The first damped-Jacobi sweep from zero, x1 = w b / diag(A).
template <typename T> struct VarInSmooth0 {
    const T* b;
    double omega;
    static constexpr bool kDiag = true, kCol = false;
    using Raw = double;
    __device__ __forceinline__ Raw fetch(int i, int j, int N) const { return (double)b[gidx(i, j, N)]; }
    __device__ __forceinline__ double make(const Raw& r, int, int, double dinv) const { return omega * r * dinv; }
};

// x + P e_c, P the bilinear prolongation from the coarse side Nc = (N+1)/2.
template <typename T> struct VarInProlong {
    const T* x;
    const T* ec;
    int Nc;
    static constexpr bool kDiag = false, kCol = true;
    struct Raw { double x, e00, e10, e01, e11; };
    __device__ __forceinline__ Raw fetch(int i, int j, int N) const {
        const int ic = i >> 1, jc = j >> 1;
        const int ic1 = ic + 1 < Nc ? ic + 1 : ic, jc1 = jc + 1 < Nc ? jc + 1 : jc;
        Raw r;
        r.x = (double)x[gidx(i, j, N)];
        r.e00 = (double)ec[gidx(ic, jc, Nc)];
        r.e10 = (double)ec[gidx(ic1, jc, Nc)];
        r.e01 = (double)ec[gidx(ic, jc1, Nc)];
        r.e11 = (double)ec[gidx(ic1, jc1, Nc)];
        return r;
    }
    __device__ __forceinline__ double make(const Raw& r, int i, int j, double) const {
        double v;
        if (!(i & 1)) v = (!(j & 1)) ? r.e00 : (r.e00 + r.e01) * 0.5;
        else v = (!(j & 1)) ? (r.e00 + r.e10) * 0.5 : (r.e00 + r.e10 + r.e01 + r.e11) * 0.25;
        return r.x + v;
    }
    // w[k] = (x + P e_c)(i, jbeg + k), k < W, jbeg of parity PAR: each coarse
    // row loaded once.  Rows jbeg .. jbeg+W-1 lie inside the grid and i < N
    // (an odd i has i/2 + 1 < Nc).
    template <int W, int PAR>
    __device__ __forceinline__ void load_col(int i, int jbeg, int N, double* w) const {
        constexpr int NC = ((PAR + W - 1) >> 1) + 2;
        const int ic = i >> 1, jc0 = jbeg >> 1;
        const bool odd = i & 1;
        double c[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            const T* e = ec + gidx(ic, jc0 + m, Nc);
            c[m] = odd ? ((double)e[0] + (double)e[1]) * 0.5 : (double)e[0];
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int s = PAR + k;
            const double v = (s & 1) ? (c[s >> 1] + c[(s >> 1) + 1]) * 0.5 : c[s >> 1];
            w[k] = (double)x[gidx(i, jbeg + k, N)] + v;
        }
    }
};

// out = s A u
template <typename T> struct VarOutStore {
    T* o;
    double s;
    static constexpr bool kDiag = false, kDot = false;
    using Aux = double;
    __device__ __forceinline__ Aux pre(size_t) const { return 0.0; }
    __device__ __forceinline__ void emit(size_t q, double, double Au, double, double, double&) const {
        o[q] = (T)(s * Au);
    }
};

// out = A u, partial += a' out
template <typename T> struct VarOutStoreDot {
    T* o;
    const T* a;
    static constexpr bool kDiag = false, kDot = true;
    using Aux = double;
    __device__ __forceinline__ Aux pre(size_t q) const { return (double)a[q]; }
    __device__ __forceinline__ void emit(size_t q, double, double Au, double, double av, double& acc) const {
        o[q] = (T)Au;
        acc += av * Au;
    }
};

// r = b - A u, partial += r'r
template <typename T> struct VarOutResid {
    const T* b;
    T* r;
    static constexpr bool kDiag = false, kDot = true;
    using Aux = double;
    __device__ __forceinline__ Aux pre(size_t q) const { return (double)b[q]; }
    __device__ __forceinline__ void emit(size_t q, double, double Au, double, double bq, double& acc) const {
        const double rv = bq - Au;
        r[q] = (T)rv;
        acc += rv * rv;
    }
};

// Damped-Jacobi sweep xo = u + w (b - A u) / diag(A).
template <typename T> struct VarOutSmooth {
    const T* b;
    T* xo;
    double omega;
    static constexpr bool kDiag = true, kDot = false;
    using Aux = double;
    __device__ __forceinline__ Aux pre(size_t q) const { return (double)b[q]; }
    __device__ __forceinline__ void emit(size_t q, double u, double Au, double dinv, double bq, double&) const {
        xo[q] = (T)(u + omega * (bq - Au) * dinv);
    }
};

// The same sweep with partial += b' xo (r'z after a V-cycle's last sweep).
template <typename T> struct VarOutSmoothDot {
    const T* b;
    T* xo;
    double omega;
    static constexpr bool kDiag = true, kDot = true;
    using Aux = double;
    __device__ __forceinline__ Aux pre(size_t q) const { return (double)b[q]; }
    __device__ __forceinline__ void emit(size_t q, double u, double Au, double dinv, double bq, double& acc) const {
        const double o = u + omega * (bq - Au) * dinv;
        xo[q] = (T)o;
        acc += bq * o;
    }
};

// The fused Krylov step on the search direction (FP64): the operand is
// p_new = z + beta p_old (beta from the device scalar block); the output writes
// p_new and A p_new, applies the pending x += alpha p_old of the previous
// iteration and accumulates p_new' A p_new.
struct VarInAxpy {
    const double* r;
    const double* p;
    const double* beta;
    double bt;
    static constexpr bool kDiag = false, kCol = false;
    struct Raw { double r, p; };
    __device__ __forceinline__ void init() { bt = *beta; }
    __device__ __forceinline__ Raw fetch(int i, int j, int N) const {
        const size_t q = gidx(i, j, N);
        Raw v;
        v.r = r[q];
        v.p = p[q];
        return v;
    }
    __device__ __forceinline__ double make(const Raw& v, int, int, double) const { return v.r + bt * v.p; }
};
struct VarOutPUpd {
    double* pnew;
    double* Ap;
    double* x;
    const double* pold;
    const double* alpha;
    double al;
    static constexpr bool kDiag = false, kDot = true;
    struct Aux { double x, p; };
    __device__ __forceinline__ void init() { al = *alpha; }
    __device__ __forceinline__ Aux pre(size_t q) const {
        Aux a;
        a.x = x[q];
        a.p = pold[q];
        return a;
    }
    __device__ __forceinline__ void emit(size_t q, double u, double Au, double, const Aux& a, double& acc) const {
        pnew[q] = u;
        Ap[q] = Au;
        x[q] = a.x + al * a.p;
        acc += u * Au;
    }
};

// The operand at any point (face rows and closure columns): d from the table.
template <class IN>
__device__ __forceinline__ double var_operand(const IN& in, const VarTab& t, const double* kb, const double* mb,
                                              int N, int i, int j) {
    return in.make(in.fetch(i, j, N), i, j, IN::kDiag ? 1.0 / var_diag(t, kb, mb, N, i, j) : 0.0);
}

// Tiled evaluation for blocks whose rows (with halo) avoid the closure rows:
// each thread keeps its column window of the operand, rows j0-H ..
// j0+kVarRows-1+H, in registers for the j-direction; an LDS tile of the
// operand on the block's rows serves the i-direction (centre column of
// thread tid at tid + HL); the i-band coefficients are loaded once per thread
// and the j-band ones of the block's rows go to LDS (kr) with the operand
// (read after the output stores have begun they could not come through the
// scalar cache: the compiler cannot rule out that the stores reach them);
// H weights along j are 1 there.  With ICHK (the first and last columns of
// blocks) each thread tests its column: out-of-grid threads idle, and
// threads on the closure columns take the closure row of m (table t staged
// in LDS) and their H weight.  The tile's left halo is then HL = C columns
// wide, so every closure column a thread of the block reads is in the tile:
// at the left face they are the first block's own columns, at the right face
// the last C columns, which a last block narrower than C reaches through
// that halo.
template <int P> constexpr int var_tile_hl() { return VarShape<P>::C; }
template <int P> constexpr int var_tile_w() { return kVarThreads + var_tile_hl<P>() + P / 2; }

template <int P, bool ICHK, typename T, class IN, class OUT>
__device__ __forceinline__ double var_tile(const IN& in, const OUT& out, const VarTab* __restrict__ tp,
                                           const double* __restrict__ kb, const double* __restrict__ mb,
                                           const VarTab& t, int i0, int j0, double (*tile)[var_tile_w<P>()],
                                           double (*kr)[kVarRows]) {
    constexpr int H = P / 2;
    constexpr int J = kVarRows;
    constexpr int W = J + 2 * H;
    constexpr int HL = var_tile_hl<P>();
    constexpr int C = VarShape<P>::C;
    constexpr int NL = ICHK ? HL : H;          // left halo columns loaded
    constexpr bool kD = IN::kDiag || OUT::kDiag;
    const int N = tp->N, R = tp->R;
    const int tid = threadIdx.x;
    const int i = i0 + tid;
    const bool act = !ICHK || i < N;
    const bool iint = !ICHK || (i >= R && i < N - R);
    double w[W];
    // loads first: the operand window, the output's pre-loads, the i-band
    // coefficients of this column
    typename IN::Raw raw[IN::kCol ? 1 : W];
    if (act) {
        if constexpr (IN::kCol) {
            in.template load_col<W, (H & 1)>(i, j0 - H, N, w);   // j0 is a multiple of the even J
        } else {
#pragma unroll
            for (int k = 0; k < W; ++k) raw[k] = in.fetch(i, j0 - H + k, N);
        }
    }
    typename OUT::Aux aux[J];
#pragma unroll
    for (int r = 0; r < J; ++r) aux[r] = act ? out.pre(gidx(i, j0 + r, N)) : typename OUT::Aux{};
    double cm[2 * H + 1];
#pragma unroll
    for (int q = 0; q <= 2 * H; ++q) cm[q] = iint ? mb[(size_t)q * N + i] : 0.0;
    if (tid < (2 * H + 1) * J) {
        const int q = tid / J, r = tid - q * J;
        kr[q][r] = kb[(size_t)q * N + j0 + r];
    }
    // halo columns: NL left, H right, J rows
    for (int e = tid; e < (NL + H) * J; e += kVarThreads) {
        const int r = e / (NL + H), c = e - r * (NL + H);
        const int ii = c < NL ? i0 - NL + c : i0 + kVarThreads + (c - NL);
        double v = 0.0;
        if (!ICHK || (ii >= 0 && ii < N)) {
            if constexpr (ICHK) v = var_operand(in, t, kb, mb, N, ii, j0 + r);
            else v = in.make(in.fetch(ii, j0 + r, N), ii, j0 + r,
                             IN::kDiag ? 1.0 / (kb[(size_t)H * N + j0 + r] + mb[(size_t)H * N + ii]) : 0.0);
        }
        tile[r][c < NL ? HL - NL + c : HL + kVarThreads + (c - NL)] = v;
    }
    double hwi = 1.0, mdi = 0.0;
    if (ICHK && act) {
        hwi = var_hw(t, N, i);
        if (kD && !iint) mdi = var_diag1(t.m, mb, R, H, N, i);
    }
    // 1 / diag(A) on the window's rows for an operand that needs it (the
    // first sweep from zero; the output sweep reuses the centre rows); an
    // output alone forms it per row after the barrier.  hw_j = 1 on these rows.
    constexpr int WD = IN::kDiag ? W : 1;
    double dinv[WD];
    if constexpr (IN::kDiag) {
#pragma unroll
        for (int k = 0; k < W; ++k)
            dinv[k] = act ? 1.0 / (hwi * kb[(size_t)H * N + j0 - H + k] + (iint ? cm[H] : mdi)) : 0.0;
    }
    if constexpr (!IN::kCol) {
#pragma unroll
        for (int k = 0; k < W; ++k) w[k] = act ? in.make(raw[k], i, j0 - H + k, IN::kDiag ? dinv[k] : 0.0) : 0.0;
    } else if (!act) {
#pragma unroll
        for (int k = 0; k < W; ++k) w[k] = 0.0;
    }
#pragma unroll
    for (int r = 0; r < J; ++r) tile[r][tid + HL] = w[r + H];
    __syncthreads();
    double acc = 0.0;
    if (!act) return acc;
#pragma unroll
    for (int r = 0; r < J; ++r) {
        const int j = j0 + r;
        double kp = 0.0, mp = 0.0;
#pragma unroll
        for (int q = 0; q <= 2 * H; ++q) kp += kr[q][r] * w[r + q];
        if (iint) {
#pragma unroll
            for (int q = 0; q <= 2 * H; ++q) mp += cm[q] * tile[r][tid + HL - H + q];
        } else {
            // closure row of m from the tile (the order of var_row); a column
            // outside it (the block before a last block wider than H, only
            // for N not of the form 2^k + 1) from memory
            const bool lo = i < R;
            const int rr = lo ? i : N - 1 - i;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int cc = lo ? c : N - 1 - c;
                const double v = cc < i0 + kVarThreads + H ? tile[r][cc - i0 + HL] : var_operand(in, t, kb, mb, N, cc, j);
                mp += (lo ? t.m.lo[rr][c] : t.m.hi[rr][c]) * v;
            }
        }
        double di = 0.0;
        if constexpr (OUT::kDiag) di = IN::kDiag ? dinv[r + H] : 1.0 / (hwi * kr[H][r] + (iint ? cm[H] : mdi));
        out.emit(gidx(i, j, N), w[r + H], hwi * kp + mp, di, aux[r], acc);
    }
    return acc;
}

// One pass of the variable-coefficient operator over the N x N grid with
// operand IN and output OUT (one problem; a launch on a converged solve,
// sc->done, returns at once).  kb / mb: the table's coefficient bands (as
// kernel arguments their loads are known to be global-memory loads).  Blocks: kVarThreads x kVarRows tiles, nbx
// along i, the face tiles first and the interior ones remapped so each XCD
// walks a contiguous band (tile_order).
template <int P, typename T, class IN, class OUT>
__global__ __launch_bounds__(kVarThreads) void var_pass(IN in, OUT out, const VarTab* __restrict__ tp,
                                                       const double* __restrict__ kb, const double* __restrict__ mb,
                                                       const Scalars* sc, double* partial, int nbx, int nby) {
    constexpr int H = P / 2;
    constexpr int J = kVarRows;
    if (sc && sc->done) return;
    __shared__ VarTab t;
    __shared__ double tile[J][var_tile_w<P>()];
    __shared__ double kr[2 * H + 1][J];
    const int N = tp->N, R = tp->R;
    const int tid = threadIdx.x;
    int bx, by;
    {   // face tiles first, then the interior ones in per-XCD bands (tile_order)
        const int ylast = N - R - H - J, xlast = N - R - H - kVarThreads;
        int by0 = (R + H + J - 1) / J, by1 = ylast >= 0 ? min(nby, ylast / J + 1) : 0;
        int bx0 = (R + H + kVarThreads - 1) / kVarThreads, bx1 = xlast >= 0 ? min(nbx, xlast / kVarThreads + 1) : 0;
        if (by1 <= by0 || bx1 <= bx0) bx0 = bx1 = by0 = by1 = 0;
        tile_order(nbx, nby, bx0, bx1, by0, by1, (int)blockIdx.x, true, bx, by);
    }
    const int blk = by * nbx + bx;
    const int i0 = bx * kVarThreads, j0 = by * J;
    if constexpr (has_init<IN>::value) in.init();
    if constexpr (has_init<OUT>::value) out.init();
    const bool jin = j0 - H >= R && j0 + J + H <= N - R;
    const bool iin = i0 - H >= R && i0 + kVarThreads + H <= N - R;
    double acc = 0.0;
    if (jin && iin) {
        acc = var_tile<P, false, T>(in, out, tp, kb, mb, t, i0, j0, tile, kr);
    } else {
        {
            const unsigned long long* src = reinterpret_cast<const unsigned long long*>(tp);
            unsigned long long* dst = reinterpret_cast<unsigned long long*>(&t);
            for (int k = tid; k < (int)(sizeof(VarTab) / 8); k += kVarThreads) dst[k] = src[k];
        }
        __syncthreads();
        if (jin) {
            acc = var_tile<P, true, T>(in, out, tp, kb, mb, t, i0, j0, tile, kr);
        } else {
            const int i = i0 + tid;   // face rows: per point
            if (i < N) {
                const double hwi = var_hw(t, N, i);
                for (int r = 0; r < J; ++r) {
                    const int j = j0 + r;
                    if (j >= N) break;
                    const double hwj = var_hw(t, N, j);
                    const double kp = var_row<H>(t.k, kb, t.R, t.C, N, j, [&](int c) { return var_operand(in, t, kb, mb, N, i, c); });
                    const double mp = var_row<H>(t.m, mb, t.R, t.C, N, i, [&](int c) { return var_operand(in, t, kb, mb, N, c, j); });
                    const double dinv = OUT::kDiag ? 1.0 / var_diag(t, kb, mb, N, i, j) : 0.0;
                    const size_t q = gidx(i, j, N);
                    out.emit(q, var_operand(in, t, kb, mb, N, i, j), hwi * kp + hwj * mp, dinv, out.pre(q), acc);
                }
            }
        }
    }
    if constexpr (OUT::kDot) {
        __syncthreads();
        const double s = block_sum<double, kVarThreads>(acc);
        if (threadIdx.x == 0) partial[blk] = s;
    }
}

// rc = scale * FW(b - A u) on the coarse grid of side Nc = (N+1)/2, the fine
// residual kept on chip: a block owns kVarThreads/2 - 1 coarse columns and
// kVarRows/2 coarse rows and forms the residual on the kVarThreads x
// (kVarRows + 1) fine points they weigh (the geometry of resid_restrict);
// interior rows through the register window and LDS tile, face rows per
// point.  Full weighting as resid_restrict.
template <int P, typename T>
__global__ __launch_bounds__(kVarThreads) void var_resid_restrict(const T* __restrict__ u, const T* __restrict__ b,
                                                                 T* __restrict__ rc, double scale,
                                                                 const VarTab* __restrict__ tp,
                                                                 const double* __restrict__ kb,
                                                                 const double* __restrict__ mb, int Nc, int nbx,
                                                                 int nby, const Scalars* sc) {
    constexpr int H = P / 2;
    constexpr int TI = kVarThreads;
    constexpr int JR = kVarRows + 1;      // fine residual rows per block
    constexpr int W = JR + 2 * H;         // column window
    constexpr int CI = TI / 2 - 1;        // coarse columns per block
    constexpr int CJ = kVarRows / 2;      // coarse rows per block
    if (sc && sc->done) return;
    __shared__ VarTab t;
    __shared__ double ut[JR][TI + 2 * H];
    __shared__ double kr[2 * H + 1][JR];  // the j-band coefficients of the block's rows
    double (*rt)[TI + 2 * H] = ut;        // the residual reuses the operand tile after a barrier
    const int N = tp->N, R = tp->R;
    const int tid = threadIdx.x;
    int bx, by;
    {   // face blocks first (tile_order); interior blocks: the fine points they load lie off the closure rows
        const int ylast = N - R - H - kVarRows, xlast = N - R - H - TI + 1;
        int by0 = (R + H + 1 + kVarRows - 1) / kVarRows, by1 = ylast >= 0 ? min(nby, ylast / kVarRows + 1) : 0;
        int bx0 = (R + H + 1 + 2 * CI - 1) / (2 * CI), bx1 = xlast >= 0 ? min(nbx, xlast / (2 * CI) + 1) : 0;
        if (by1 <= by0 || bx1 <= bx0) bx0 = bx1 = by0 = by1 = 0;
        tile_order(nbx, nby, bx0, bx1, by0, by1, (int)blockIdx.x, true, bx, by);
    }
    const int Ic0 = bx * CI, Jc0 = by * CJ;
    const int fi0 = 2 * Ic0 - 1, fj0 = 2 * Jc0 - 1;   // fine origin (may be -1)
    const int fi = fi0 + tid;
    const bool jin = fj0 - H >= R && fj0 + JR + H <= N - R;
    const bool iin = fi0 - H >= R && fi0 + TI + H <= N - R;
    if (!(jin && iin)) {
        const unsigned long long* src = reinterpret_cast<const unsigned long long*>(tp);
        unsigned long long* dst = reinterpret_cast<unsigned long long*>(&t);
        for (int k = tid; k < (int)(sizeof(VarTab) / 8); k += TI) dst[k] = src[k];
        __syncthreads();
    }
    double res[JR];
    if (jin) {
        const bool act = iin || (fi >= 0 && fi < N);
        const bool iint = iin || (fi >= R && fi < N - R);
        double w[W], cm[2 * H + 1], bq[JR];
#pragma unroll
        for (int k = 0; k < W; ++k) w[k] = act ? (double)u[gidx(fi, fj0 - H + k, N)] : 0.0;
#pragma unroll
        for (int q = 0; q <= 2 * H; ++q) cm[q] = (act && iint) ? mb[(size_t)q * N + fi] : 0.0;
#pragma unroll
        for (int e0 = 0; e0 < 2 * H * JR; e0 += TI) {
            const int e = e0 + tid;
            if (e < 2 * H * JR) {
                const int side = e / (H * JR), rem = e % (H * JR);
                const int off = rem / JR, r = rem % JR;
                const int ii = side == 0 ? fi0 - H + off : fi0 + TI + off;
                ut[r][side == 0 ? off : TI + H + off] =
                    (iin || (ii >= 0 && ii < N)) ? (double)u[gidx(ii, fj0 + r, N)] : 0.0;
            }
        }
        if (tid < (2 * H + 1) * JR) {
            const int q = tid / JR, r = tid - q * JR;
            kr[q][r] = kb[(size_t)q * N + fj0 + r];
        }
#pragma unroll
        for (int r = 0; r < JR; ++r) {
            ut[r][tid + H] = w[r + H];
            bq[r] = act ? (double)b[gidx(fi, fj0 + r, N)] : 0.0;
        }
        __syncthreads();
        const double hwi = iin ? 1.0 : (act ? var_hw(t, N, fi) : 0.0);
#pragma unroll
        for (int r = 0; r < JR; ++r) {
            const int j = fj0 + r;
            double kp = 0.0, mp = 0.0;
#pragma unroll
            for (int q = 0; q <= 2 * H; ++q) kp += kr[q][r] * w[r + q];
            if (iint) {
#pragma unroll
                for (int q = 0; q <= 2 * H; ++q) mp += cm[q] * ut[r][tid + q];
            } else if (act) {
                mp = var_row<H>(t.m, mb, R, t.C, N, fi, [&](int c) { return (double)u[gidx(c, j, N)]; });
            }
            res[r] = act ? bq[r] - (hwi * kp + mp) : 0.0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < JR; ++r) {
            const int fj = fj0 + r;
            double v = 0.0;
            if (fi >= 0 && fi < N && fj >= 0 && fj < N) {
                const double kp = var_row<H>(t.k, kb, R, t.C, N, fj, [&](int c) { return (double)u[gidx(fi, c, N)]; });
                const double mp = var_row<H>(t.m, mb, R, t.C, N, fi, [&](int c) { return (double)u[gidx(c, fj, N)]; });
                v = (double)b[gidx(fi, fj, N)] - (var_hw(t, N, fi) * kp + var_hw(t, N, fj) * mp);
            }
            res[r] = v;
        }
    }
    __syncthreads();                      // every thread is done reading the operand tile
#pragma unroll
    for (int r = 0; r < JR; ++r) rt[r][tid] = res[r];
    __syncthreads();
    for (int o = tid; o < CI * CJ; o += TI) {
        const int cI = o % CI, cJ = o / CI;
        const int I = Ic0 + cI, Jc = Jc0 + cJ;
        if (I >= Nc || Jc >= Nc) continue;
        const int lc = 2 * cI + 1, lr = 2 * cJ + 1;
        const double s0 = rt[lr][lc] + 0.5 * (rt[lr][lc - 1] + rt[lr][lc + 1]);
        const double sm = rt[lr - 1][lc] + 0.5 * (rt[lr - 1][lc - 1] + rt[lr - 1][lc + 1]);
        const double sp = rt[lr + 1][lc] + 0.5 * (rt[lr + 1][lc - 1] + rt[lr + 1][lc + 1]);
        rc[gidx(I, Jc, Nc)] = (T)(scale * 0.25 * (s0 + 0.5 * (sm + sp)));
    }
}

// First damped-Jacobi sweep from zero: x = omega b / diag(A), the diagonal
// hw_i k_jj + hw_j m_ii formed from the tables (blockIdx.y = row j).
template <typename T>
__global__ __launch_bounds__(256) void var_jacobi0(T* __restrict__ x, const T* __restrict__ b,
                                                  const VarTab* __restrict__ tp, double omega, const Scalars* sc) {
    if (sc && sc->done) return;
    const VarTab& t = *tp;
    const int N = t.N;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= N) return;
    const int ri = i < N - 1 - i ? i : N - 1 - i, rj = j < N - 1 - j ? j : N - 1 - j;
    const double hwi = ri < t.RH ? t.hw[ri] : 1.0, hwj = rj < t.RH ? t.hw[rj] : 1.0;
    const double d = hwi * var_diag1(t.k, t.k.band, t.R, t.H, N, j) + hwj * var_diag1(t.m, t.m.band, t.R, t.H, N, i);
    const size_t q = gidx(i, j, N);
    x[q] = (T)(omega * (double)b[q] / d);
}

// Dense A (row-major, n = N*N) of the variable-coefficient operator from
// the host tables and host bands (coarsest-level factorisation).
inline void dense_operator_var(const VarTab& t, const double* kband, const double* mband, std::vector<double>& A) {
    const int N = t.N, n = N * N, H = t.H;
    A.assign((size_t)n * n, 0.0);
    auto hwf = [&](int r) {
        const int rr = r < N - 1 - r ? r : N - 1 - r;
        return rr < t.RH ? t.hw[rr] : 1.0;
    };
    auto coef = [&](const VarAxis& a, const double* band, int r, int c) -> double {
        if (r < t.R) return c < t.C ? a.lo[r][c] : 0.0;
        if (r >= N - t.R) {
            const int rr = N - 1 - r, cc = N - 1 - c;
            return (cc >= 0 && cc < t.C) ? a.hi[rr][cc] : 0.0;
        }
        const int s = c - r;
        return (s >= -H && s <= H) ? band[(size_t)(s + H) * N + r] : 0.0;
    };
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t row = (size_t)(i + j * N) * n;
            for (int c = 0; c < N; ++c) {
                const double kv = coef(t.k, kband, j, c);
                if (kv != 0.0) A[row + (i + c * N)] += hwf(i) * kv;
                const double mv = coef(t.m, mband, i, c);
                if (mv != 0.0) A[row + (c + j * N)] += hwf(j) * mv;
            }
        }
}

}  // namespace sbp

"""CPU oracle for the SBP-SAT 2-D Poisson path (test infrastructure only).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
The product path (``sc_2022_amd``) never routes through it.

It restates, with numpy/scipy.sparse, the reference's CPU construction of the
operator, literally (same kron products, same SAT terms):

* ``diagonal_sbp_D1`` -- diagonal_sbp.jl:56-147 (p = 2, 4, 6, 8)
* ``diagonal_sbp_D2`` -- diagonal_sbp.jl:189-454 (p = 2 :191-197, p = 4
  :198-208, p = 6 :210-222, p = 8 :224-311, H/HI :421-428, S0/SN :446-448)
* ``variable_diagonal_sbp_D2`` -- diagonal_sbp.jl:470-740 (p = 2, 4, 6), and
  ``assemble_operator_var`` / ``assemble_rhs_var`` -- its 2-D use with one
  coefficient profile per axis (legacy_functions.jl:100-174, generalised)
* ``operators_2d`` / ``assembling_matrix`` -- level_2_multigrid_new.jl:48-188
* ``prolongation_matrix`` / ``restriction_matrix_normal`` /
  ``prolongation_2d`` / ``restriction_2d`` -- level_2_multigrid_new.jl:190-246
* ``cg_cpu`` -- CG_CPU, level_2_multigrid_new.jl:258-291
* ``jacobi_sweeps`` -- jacobi_brittany!, level_2_multigrid_new.jl:335-344
* ``modified_richardson`` -- level_2_multigrid_new.jl:346-350
* ``reference_gpu_apply_p2`` -- matrix_free_A_full_GPU and its five kernels,
  split_matrix_free.jl:32-49, 324-429, 444-486 (returns -A u)
* ``vcycle`` / ``mgpcg`` -- the multi-level V-cycle and MG-preconditioned CG
  (mg_preconditioned_CG, level_2_multigrid_new.jl:480-524, generalised from two
  levels to a hierarchy down to 17x17 with damped-Jacobi smoothing).
* ``initial_guess_interpolation_cg`` -- initial_guess_interpolation_CG and its
  three-level form, initial_guess_cg.jl:394-399, 447-456.
* ``two_level_richardson`` / ``matrix_free_mgcg`` -- the reference's own
  preconditioner and MGCG, initial_guess_cg.jl:147-180, 303-346.

Grid convention (the reference's): level k gives N = 2^k + 1 points per side
on [0,1]^2, h = 1/2^k.  A grid function is a flat vector with element (i, j)
at ``i + j*N`` (0-based); the fast index i is the operator's "y" (Neumann
S/N faces at i = 0 and i = N-1), the slow index j is the operator's "x"
(Dirichlet W/E faces at j = 0 and j = N-1).

Pinning (each function is checked in tests/test_oracle.py against an
independent fact): D1/D2 exactness on polynomials and the SBP property,
sum(H) = 1, A symmetric positive definite, the literal restatement of the
reference GPU kernel equal to -A, the closed form A = K (x) H + H (x) M,
transfer operators against their closed forms (R = P^T / 4, P exact on
bilinears), CG and MG-PCG against the sparse direct solve, the
manufactured-solution convergence rates (2 for p = 2, about 4 for p = 4), and
the interpolated initial guess against a bilinear solution, which every level
represents exactly (so the guess must equal it to round-off).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------
# 1-D operators (diagonal_sbp.jl)
# ---------------------------------------------------------------------------

_D1_TABLES = {
    # diagonal_sbp.jl:58-68
    2: dict(bhinv=[2.0], d=[-1 / 2, 0, 1 / 2], bd=[[-1.0, 1.0]]),
    4: dict(
        bhinv=[48 / 17, 48 / 59, 48 / 43, 48 / 49],
        d=[1 / 12, -2 / 3, 0, 2 / 3, -1 / 12],
        bd=[[-24 / 17, 59 / 34, -4 / 17, -3 / 34, 0, 0],
            [-1 / 2, 0, 1 / 2, 0, 0, 0],
            [4 / 43, -59 / 86, 0, 59 / 86, -4 / 43, 0],
            [3 / 98, 0, -59 / 98, 0, 32 / 49, -4 / 49]],
    ),
    6: dict(  # diagonal_sbp.jl:69-79
        bhinv=[43200 / 13649, 8640 / 12013, 4320 / 2711, 4320 / 5359, 8640 / 7877, 43200 / 43801],
        d=[-1 / 60, 3 / 20, -3 / 4, 0.0, 3 / 4, -3 / 20, 1 / 60],
        bd=[
            [-21600 / 13649, 104009 / 54596, 30443 / 81894, -33311 / 27298, 16863 / 27298, -15025 / 163788, 0.0, 0.0, 0.0],
            [-104009 / 240260, 0.0, -311 / 72078, 20229 / 24026, -24337 / 48052, 36661 / 360390, 0.0, 0.0, 0.0],
            [-30443 / 162660, 311 / 32532, 0.0, -11155 / 16266, 41287 / 32532, -21999 / 54220, 0.0, 0.0, 0.0],
            [33311 / 107180, -20229 / 21436, 485 / 1398, 0.0, 4147 / 21436, 25427 / 321540, 72 / 5359, 0.0, 0.0],
            [-16863 / 78770, 24337 / 31508, -41287 / 47262, -4147 / 15754, 0.0, 342523 / 472620, -1296 / 7877, 144 / 7877, 0.0],
            [15025 / 525612, -36661 / 262806, 21999 / 87602, -25427 / 262806, -342523 / 525612, 0.0, 32400 / 43801, -6480 / 43801, 720 / 43801],
        ],
    ),
    8: dict(  # diagonal_sbp.jl:80-91
        bhinv=[5080320 / 1498139, 725760 / 1107307, 80640 / 20761, 725760 / 1304999, 725760 / 299527, 80640 / 103097, 725760 / 670091, 5080320 / 5127739],
        d=[1 / 280, -4 / 105, 1 / 5, -4 / 5, 0.0, 4 / 5, -1 / 5, 4 / 105, -1 / 280],
        bd=[
            [-2540160 / 1498139, 699846290793 / 311403172540, -10531586157 / 311403172540, -145951651079 / 186841903524, 398124597 / 15570158627, 39152001 / 113858564, -80631892229 / 934209517620, -6230212503 / 311403172540, 0.0, 0.0, 0.0, 0.0],
            [-24132630717 / 55557028660, 0.0, 2113176981 / 23016483302, 5686186719 / 11508241651, -3408473341 / 138098899812, -39291999 / 210388330, 607046586 / 11508241651, 3460467023 / 483346149342, 0.0, 0.0, 0.0, 0.0],
            [3510528719 / 90623010660, -704392327 / 1294614438, 0.0, 503511235 / 1294614438, 354781619 / 2589228876, 407439 / 3944590, -2986842 / 16597621, 169381493 / 3020767022, 0.0, 0.0, 0.0, 0.0],
            [145951651079 / 1139279786988, -5686186719 / 13562854607, -1510533705 / 27125709214, 0.0, 6763379967 / 54251418428, 13948923 / 49589962, -1603900430 / 40688563821, -3742312557 / 189879964498, 0.0, 0.0, 0.0, 0.0],
            [-398124597 / 21790888777, 3408473341 / 37355809332, -1064344857 / 12451936444, -6763379967 / 12451936444, 0.0, 763665 / 1198108, -1282435899 / 12451936444, 7822226819 / 261490665324, -2592 / 299527, 0.0, 0.0, 0.0],
            [-1864381 / 23506116, 13097333 / 58765290, -407439 / 19588430, -4649641 / 11753058, -254555 / 1237164, 0.0, 5346432 / 9794215, -923328 / 9794215, 3072 / 103097, -288 / 103097, 0.0, 0.0],
            [11518841747 / 417855345780, -607046586 / 6964255763, 1192698 / 23768791, 1603900430 / 20892767289, 1282435899 / 27857023052, -48117888 / 63658645, 0.0, 301190400 / 366539777, -145152 / 670091, 27648 / 670091, -2592 / 670091, 0.0],
            [6230212503 / 1065851828540, -3460467023 / 319755548562, -1524433437 / 106585182854, 3742312557 / 106585182854, -7822226819 / 639511097124, 58169664 / 487135205, -2108332800 / 2804873233, 0.0, 4064256 / 5127739, -1016064 / 5127739, 193536 / 5127739, -18144 / 5127739],
        ],
    ),
}

_D2_TABLES = {
    # diagonal_sbp.jl:191-208
    2: dict(bhinv=[2.0], d=[1.0, -2.0, 1.0], bd=[[1.0, -2.0, 1.0]],
            BS=[3 / 2, -2.0, 1 / 2]),
    4: dict(
        bhinv=[48 / 17, 48 / 59, 48 / 43, 48 / 49],
        d=[-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12],
        bd=[[2, -5, 4, -1, 0, 0],
            [1, -2, 1, 0, 0, 0],
            [-4 / 43, 59 / 43, -110 / 43, 59 / 43, -4 / 43, 0],
            [-1 / 49, 0, 59 / 49, -118 / 49, 64 / 49, -4 / 49]],
        BS=[11 / 6, -3.0, 3 / 2, -1 / 3],
    ),
    6: dict(  # diagonal_sbp.jl:210-222
        bhinv=[43200 / 13649, 8640 / 12013, 4320 / 2711, 4320 / 5359, 8640 / 7877, 43200 / 43801],
        d=[1 / 90, -3 / 20, 3 / 2, -49 / 18, 3 / 2, -3 / 20, 1 / 90],
        bd=[
            [114170 / 40947, -438107 / 54596, 336409 / 40947, -276997 / 81894, 3747 / 13649, 21035 / 163788, 0.0, 0.0, 0.0],
            [6173 / 5860, -2066 / 879, 3283 / 1758, -303 / 293, 2111 / 3516, -601 / 4395, 0.0, 0.0, 0.0],
            [-52391 / 81330, 134603 / 32532, -21982 / 2711, 112915 / 16266, -46969 / 16266, 30409 / 54220, 0.0, 0.0, 0.0],
            [68603 / 321540, -12423 / 10718, 112915 / 32154, -75934 / 16077, 53369 / 21436, -54899 / 160770, 48 / 5359, 0.0, 0.0],
            [-7053 / 39385, 86551 / 94524, -46969 / 23631, 53369 / 15754, -87904 / 23631, 820271 / 472620, -1296 / 7877, 96 / 7877, 0.0],
            [21035 / 525612, -24641 / 131403, 30409 / 87602, -54899 / 131403, 820271 / 525612, -117600 / 43801, 64800 / 43801, -6480 / 43801, 480 / 43801],
        ],
        BS=[25 / 12, -4.0, 3.0, -4 / 3, 1 / 4],
    ),
    8: dict(  # diagonal_sbp.jl:224-311
        bhinv=[5080320 / 1498139, 725760 / 1107307, 80640 / 20761, 725760 / 1304999, 725760 / 299527, 80640 / 103097, 725760 / 670091, 5080320 / 5127739],
        d=[-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
        bd=[
            [4870382994799 / 1358976868290, -893640087518 / 75498714905, 926594825119 / 60398971924, -1315109406200 / 135897686829, 39126983272 / 15099742981, 12344491342 / 75498714905, -451560522577 / 2717953736580, 0.0, 0.0, 0.0, 0.0, 0.0],
            [333806012194 / 390619153855, -154646272029 / 111605472530, 1168338040 / 33481641759, 82699112501 / 133926567036, -171562838 / 11160547253, -28244698346 / 167408208795, 11904122576 / 167408208795, -2598164715 / 312495323084, 0.0, 0.0, 0.0, 0.0],
            [7838984095 / 52731029988, 1168338040 / 5649753213, -88747895 / 144865467, 423587231 / 627750357, -43205598281 / 22599012852, 4876378562 / 1883251071, -5124426509 / 3766502142, 10496900965 / 39548272491, 0.0, 0.0, 0.0, 0.0],
            [-94978241528 / 828644350023, 82699112501 / 157837019052, 1270761693 / 13153084921, -167389605005 / 118377764289, 48242560214 / 39459254763, -31673996013 / 52612339684, 43556319241 / 118377764289, -44430275135 / 552429566682, 0.0, 0.0, 0.0, 0.0],
            [1455067816 / 21132528431, -171562838 / 3018932633, -43205598281 / 36227191596, 48242560214 / 9056797899, -52276055645 / 6037865266, 57521587238 / 9056797899, -80321706377 / 36227191596, 8078087158 / 21132528431, -1296 / 299527, 0.0, 0.0, 0.0],
            [10881504334 / 327321118845, -28244698346 / 140280479505, 4876378562 / 9352031967, -10557998671 / 12469375956, 57521587238 / 28056095901, -278531401019 / 93520319670, 73790130002 / 46760159835, -137529995233 / 785570685228, 2048 / 103097, -144 / 103097, 0.0, 0.0],
            [-135555328849 / 8509847458140, 11904122576 / 101307707835, -5124426509 / 13507694378, 43556319241 / 60784624701, -80321706377 / 81046166268, 73790130002 / 33769235945, -950494905688 / 303923123505, 239073018673 / 141830790969, -145152 / 670091, 18432 / 670091, -1296 / 670091, 0.0],
            [0.0, -2598164715 / 206729925524, 10496900965 / 155047444143, -44430275135 / 310094888286, 425162482 / 2720130599, -137529995233 / 620189776572, 239073018673 / 155047444143, -144648000000 / 51682481381, 8128512 / 5127739, -1016064 / 5127739, 129024 / 5127739, -9072 / 5127739],
        ],
        BS=[4723 / 2100, -839 / 175, 157 / 35, -278 / 105, 103 / 140, 1 / 175, -6 / 175],
    ),
}


def _check_order(p):
    if p not in (2, 4, 6, 8):
        raise ValueError(f"Operators for order {p} are not implemented")


def _banded(p, N, d, bd, mirror_sign):
    """Interior stencil d plus closure block bd and its mirror image."""
    bd = np.asarray(bd, dtype=np.float64)
    bm, bn = bd.shape
    M = N + 1
    if M < 2 * bm or M < bn:
        raise ValueError(
            f"Grid not big enough to support the operator. Grid must have N >= {max(bn, 2 * bm)}")
    rows, cols, vals = [], [], []
    half = p // 2
    for r in range(bm, M - bm):                      # B_I1/B_J1/B_V1
        for k, dv in enumerate(d):
            rows.append(r); cols.append(r - half + k); vals.append(dv)
    for r in range(bm):                              # B_I2/B_J2/B_V2
        for c in range(bn):
            rows.append(r); cols.append(c); vals.append(bd[r, c])
    for r in range(bm):                              # B_I3/B_J3/B_V3
        for c in range(bn):
            rows.append(M - 1 - r); cols.append(M - 1 - c)
            vals.append(mirror_sign * bd[r, c])
    return rows, cols, vals, bm


def diagonal_sbp_D1(p, N, xc=(-1.0, 1.0)):
    """(D, HI, H, r) -- diagonal_sbp.jl:56-147 (p = 2, 4, 6, 8)."""
    _check_order(p)
    t = _D1_TABLES[p]
    rows, cols, vals, bm = _banded(p, N, t["d"], t["bd"], -1.0)
    M = N + 1
    h = (xc[1] - xc[0]) / N
    assert h > 0
    HV = np.ones(M)
    bh = np.asarray(t["bhinv"])
    HV[:bm] = 1.0 / bh
    HV[M - bm:] = 1.0 / bh[::-1]
    H = sp.diags(h * HV).tocsr()
    HI = sp.diags(1.0 / (h * HV)).tocsr()
    D = sp.csr_matrix((np.asarray(vals) / h, (rows, cols)), shape=(M, M))
    r = np.linspace(xc[0], xc[1], M)
    return D, HI, H, r


def diagonal_sbp_D2(p, N, xc=(-1.0, 1.0)):
    """(D, S0, SN, HI, H, r) -- diagonal_sbp.jl:189-454 (p = 2, 4, 6, 8)."""
    _check_order(p)
    t = _D2_TABLES[p]
    rows, cols, vals, bm = _banded(p, N, t["d"], t["bd"], +1.0)
    M = N + 1
    h = (xc[1] - xc[0]) / N
    assert h > 0
    HV = np.ones(M)
    bh = np.asarray(t["bhinv"])
    HV[:bm] = bh
    HV[M - bm:] = bh[::-1]
    HI = sp.diags(HV / h).tocsr()
    H = sp.diags(h / HV).tocsr()
    D = sp.csr_matrix((np.asarray(vals) / h ** 2, (rows, cols)), shape=(M, M))
    BS = np.asarray(t["BS"])
    nb = len(BS)
    S0 = sp.csr_matrix((-BS / h, (np.zeros(nb, int), np.arange(nb))), shape=(M, M))
    SN = sp.csr_matrix((BS / h, (np.full(nb, M - 1), M - 1 - np.arange(nb))), shape=(M, M))
    r = np.linspace(xc[0], xc[1], M)
    return D, S0, SN, HI, H, r


# ---------------------------------------------------------------------------
# Variable-coefficient D2 (diagonal_sbp.jl:470-740), orders 2, 4 and 6
# ---------------------------------------------------------------------------

# p = 4 boundary block of M(B): entry (r, c) (1-based, r <= c, symmetric) is
# sum of num/den * b_k over the listed (num, den, k), b_k = B[k] (1-based)
# (diagonal_sbp.jl:527-547).
_VAR_P4_M0 = {
    (1, 1): [(12, 17, 1), (59, 192, 2), (27010400129, 345067064608, 3), (69462376031, 2070402387648, 4)],
    (1, 2): [(-59, 68, 1), (-6025413881, 21126554976, 3), (-537416663, 7042184992, 4)],
    (1, 3): [(2, 17, 1), (-59, 192, 2), (213318005, 16049630912, 4), (2083938599, 8024815456, 3)],
    (1, 4): [(3, 68, 1), (-1244724001, 21126554976, 3), (752806667, 21126554976, 4)],
    (1, 5): [(49579087, 10149031312, 3), (-49579087, 10149031312, 4)],
    (1, 6): [(-1, 784, 4), (1, 784, 3)],
    (2, 2): [(3481, 3264, 1), (9258282831623875, 7669235228057664, 3), (236024329996203, 1278205871342944, 4)],
    (2, 3): [(-59, 408, 1), (-29294615794607, 29725717938208, 3), (-2944673881023, 29725717938208, 4)],
    (2, 4): [(-59, 1088, 1), (260297319232891, 2556411742685888, 3), (-60834186813841, 1278205871342944, 4)],
    (2, 5): [(-1328188692663, 37594290333616, 3), (1328188692663, 37594290333616, 4)],
    (2, 6): [(-8673, 2904112, 3), (8673, 2904112, 4)],
    (3, 3): [(1, 51, 1), (59, 192, 2), (13777050223300597, 26218083221499456, 4), (564461, 13384296, 5), (378288882302546512209, 270764341349677687456, 3)],
    (3, 4): [(1, 136, 1), (-125059, 743572, 5), (-4836340090442187227, 5525802884687299744, 3), (-17220493277981, 89177153814624, 4)],
    (3, 5): [(-10532412077335, 42840005263888, 4), (1613976761032884305, 7963657098519931984, 3), (564461, 4461432, 5)],
    (3, 6): [(-960119, 1280713392, 4), (-3391, 6692148, 5), (33235054191, 26452850508784, 3)],
    (4, 4): [(3, 1088, 1), (507284006600757858213, 475219048083107777984, 3), (1869103, 2230716, 5), (1, 24, 6), (1950062198436997, 3834617614028832, 4)],
    (4, 5): [(-4959271814984644613, 20965546238960637264, 3), (-1, 6, 6), (-15998714909649, 37594290333616, 4), (-375177, 743572, 5)],
    (4, 6): [(-368395, 2230716, 5), (752806667, 539854092016, 3), (1063649, 8712336, 4), (1, 8, 6)],
    (5, 5): [(8386761355510099813, 128413970713633903242, 3), (2224717261773437, 2763180339520776, 4), (5, 6, 6), (1, 24, 7), (280535, 371786, 5)],
    (5, 6): [(-35039615, 213452232, 4), (-1, 6, 7), (-13091810925, 13226425254392, 3), (-1118749, 2230716, 5), (-1, 2, 6)],
    (6, 6): [(3290636, 80044587, 4), (5580181, 6692148, 5), (5, 6, 7), (1, 24, 8), (660204843, 13226425254392, 3), (3, 4, 6)],
}


# p = 6 boundary block of M(B): entry (r, c) (1-based, r <= c, symmetric) is
# sum of w * b_k over the listed (w, k) (diagonal_sbp.jl:577-624).
_VAR_P6_M0 = {
    (1, 1): [(0.79126675946955820939, 1), (0.29684720906380007429, 2), (0.0031855190887964290152, 3), (0.016324040425909519534, 4), (0.031603022440944150877, 5), (0.031679647480161052996, 6), (0.031485777339472539205, 7)],
    (1, 2): [(-1.0166893393503381444, 1), (-0.028456273704916113690, 3), (-0.041280298383492988198, 4), (-0.13922814516201405075, 5), (-0.11957773256112017666, 6), (-0.11942677565293334109, 7)],
    (1, 3): [(0.070756429372437150463, 1), (-0.18454761060241510503, 2), (-0.043641631471118923470, 4), (0.24323679072077324609, 5), (0.15821270735372154440, 6), (0.16023485783647863076, 7)],
    (1, 4): [(0.22519915328913532127, 1), (-0.16627487110970548953, 2), (0.027105309616486712977, 3), (-0.19166461859684399091, 5), (-0.076841171601990145944, 6), (-0.082195869498316975759, 7)],
    (1, 5): [(-0.052244034642020563167, 1), (0.044400639485098762210, 2), (-0.0010239765473093878745, 3), (0.074034846453161740905, 4), (0.012416255689984968954, 6), (0.071886528478926012827, 5), (0.013793629971047355034, 7)],
    (1, 6): [(-0.018288968138771973527, 1), (0.0095746331632217580607, 2), (-0.00081057845305764042779, 3), (-0.0073488455877755196984, 4), (0.010636019497239069970, 5), (-0.013159670383826183824, 6), (-0.021179364788387535246, 7)],
    (1, 7): [(0.0019118885633161709274, 4), (-0.040681303555291499361, 5), (0.013196749810737491670, 6), (0.025572665181237836763, 7)],
    (1, 8): [(0.015596528711367857640, 5), (-0.0064861841573315378995, 6), (-0.0091103445540363197401, 7)],
    (1, 9): [(0.00055939836966298630593, 6), (-0.0013848225351007963723, 5), (0.00082542416543781006633, 7)],
    (2, 2): [(1.3063321571116676286, 1), (0.25420017604573457435, 3), (0.10438978280925626095, 4), (0.66723280210321129509, 5), (0.46818193597227494411, 6), (0.46764154101958369201, 7)],
    (2, 3): [(-0.090914102699924646049, 1), (0.11036113131714764253, 4), (-1.2903975449975188870, 5), (-0.66396052487350447871, 6), (-0.66159744640052061842, 7)],
    (2, 4): [(-0.28935573956534316666, 1), (-0.24213200040645927216, 3), (1.1876702550280310277, 5), (0.39565981499041363328, 6), (0.38600489217558000007, 7)],
    (2, 5): [(0.067127744758037639890, 1), (0.0091471926820756301800, 3), (-0.18721961430038080217, 4), (-0.13193585588531745301, 6), (-0.48715757368119118874, 5), (-0.10475163122754481381, 7)],
    (2, 6): [(0.023499279745900688694, 1), (0.0072409053835651813164, 3), (0.018583789963916794487, 4), (-0.092896161339386761743, 5), (0.12235132704188076670, 6), (0.11135203204362950339, 7)],
    (2, 7): [(-0.0048347914064469075906, 4), (0.23106838326878204031, 5), (-0.10807741421960079917, 6), (-0.11815617764273433354, 7)],
    (2, 8): [(-0.083681414344034553537, 5), (0.040934994667670546616, 6), (0.042746419676364006921, 7)],
    (2, 9): [(-0.0035765451326969831434, 6), (0.0073893991241210786821, 5), (-0.0038128539914240955387, 7)],
    (3, 3): [(0.0063271611471368738078, 1), (0.11473182007158685275, 2), (0.11667405542796800075, 4), (2.7666108082854440372, 5), (1.0709206899608171042, 6), (1.0131613910329730572, 7)],
    (3, 4): [(0.020137694138847972466, 1), (0.10337179946308864017, 2), (-2.9132216211517427243, 5), (-0.87558073434822622598, 6), (-0.69099571834888124265, 7)],
    (3, 5): [(-0.0046717510915754628683, 1), (-0.027603533656377128278, 2), (-0.19792902986208699745, 4), (0.54029853383734330523, 6), (1.2391775930319110779, 5), (0.26280380502473582273, 7)],
    (3, 6): [(-0.0016354308669218878195, 1), (-0.0059524752758832596197, 2), (0.019646827777442752194, 4), (0.32366400126390466006, 5), (-0.46595166932288709739, 6), (-0.22172727209417368594, 7)],
    (3, 7): [(-0.0051113531893524745496, 4), (-0.53558781637747543460, 5), (0.33283351044897389336, 6), (0.20786565911785401579, 7)],
    (3, 8): [(0.18243281741342895622, 5), (-0.10598160301968184459, 6), (-0.076451214393747111630, 7)],
    (3, 9): [(0.0092090899634437994856, 6), (-0.015915028188724931671, 5), (0.0067059382252811321853, 7)],
    (4, 4): [(0.064092997759871869867, 1), (0.093136576388046999489, 2), (0.23063676246347492291, 3), (3.6894403082837166203, 5), (1.1905503386876088738, 6), (0.59124795468888565194, 7)],
    (4, 5): [(-0.014868958192656041286, 1), (-0.024870405993901607642, 2), (-0.0087129289077117541871, 3), (-1.2635078373718242057, 6), (-0.30583173978439973269, 7), (-1.4706919260458029548, 5)],
    (4, 6): [(-0.0052051474298559556576, 1), (-0.0053630987475285424890, 2), (-0.0068971427657906095463, 3), (-0.78575245216674501017, 5), (0.22911480054237346001, 7), (0.99770643562927505292, 6)],
    (4, 7): [(0.66972974880676622652, 5), (-0.50132473560721279390, 6), (-0.17951612431066454373, 7)],
    (4, 8): [(-0.20229090601117515652, 5), (0.14534218580636584986, 6), (0.056948720204809306656, 7)],
    (4, 9): [(-0.012004296184410038337, 6), (-0.0047769156693859238415, 7), (0.016781211853795962179, 5)],
    (5, 5): [(0.0034494550959102336252, 1), (0.0066411834994278261016, 2), (0.00032915450832718628585, 3), (0.33577217075764772000, 4), (2.0964133295790264390, 6), (0.23173232041831268550, 7), (0.0061078257643682645765, 8), (0.71091258506833766956, 5)],
    (5, 6): [(0.0012075440723041938061, 1), (0.0014321166657521476075, 2), (0.00026055826461832559573, 3), (-0.033329411132516353908, 4), (-0.28082416973855326836, 7), (-0.027209080835250836084, 8), (0.10458654356829219874, 5), (-1.3484369866671155432, 6)],
    (5, 7): [(0.0086710380841746926251, 4), (0.17360734113554285637, 6), (0.053313621252876254126, 8), (-0.24249352624045263018, 5), (0.15690152576785882706, 7)],
    (5, 8): [(-0.086316839802171222760, 6), (0.026988423604709992435, 7), (0.080981941477156510853, 5), (-0.032764636390806391639, 8)],
    (5, 9): [(0.0074620594845308550733, 6), (-0.00081216403616686789496, 7), (0.00055227020881270902093, 8), (-0.0072021656571766961993, 5)],
    (6, 6): [(0.00042272261734493450425, 1), (0.00030882419443789644048, 2), (0.00020625757066474306202, 3), (0.0033083434042009682567, 4), (0.58280470164050018158, 5), (0.80541742203662154736, 7), (0.13383632334100334433, 8), (0.0055555555555555555556, 9), (1.1903620718618930511, 6)],
    (6, 7): [(-0.00086070442526864133026, 4), (-0.17480747086739049893, 5), (-0.31325448501150501650, 8), (-0.025000000000000000000, 9), (-0.31691663053104292713, 7), (-0.66916070916479291611, 6)],
    (6, 8): [(0.033546617916933521087, 5), (-0.33436200223869714050, 7), (0.050000000000000000000, 9), (0.21697906098076027508, 6), (0.18383632334100334433, 8)],
    (6, 9): [(0.029125184768230046430, 7), (0.022790919164749163916, 8), (-0.030689859975187405305, 6), (-0.0017817995133473605962, 5), (-0.030555555555555555556, 9)],
    (7, 7): [(0.00022392237357715991790, 4), (0.12754377854309566738, 5), (1.0116994839296081646, 6), (0.96988172751725752475, 8), (0.12500000000000000000, 9), (0.0055555555555555555556, 10), (0.48231775430312815001, 7)],
    (7, 8): [(-0.037841139730330129499, 5), (-0.29975568851348273616, 6), (-0.30000000000000000000, 9), (-0.025000000000000000000, 10), (-0.39914868674468211784, 7), (-0.43825448501150501650, 8)],
    (7, 9): [(0.046981462180226839339, 6), (-0.29668637874712374587, 8), (0.050000000000000000000, 10), (0.17163557041460064817, 7), (0.0030693461522962583624, 5), (0.17500000000000000000, 9)],
    (8, 8): [(0.012303289427168044554, 5), (0.11836475296458983325, 6), (0.94105118982279433342, 7), (0.95000000000000000000, 9), (0.12500000000000000000, 10), (0.0055555555555555555556, 11), (0.56994743445211445545, 8)],
    (8, 9): [(-0.023080678926719163396, 6), (-0.29866250537751494972, 7), (-0.30000000000000000000, 10), (-0.025000000000000000000, 11), (-0.0010477348605150508026, 5), (-0.42720908083525083608, 8), (-0.42500000000000000000, 9)],
    (9, 9): [(0.0051393702211491099770, 6), (0.12477232150094220014, 7), (0.95055227020881270902, 8), (0.95000000000000000000, 10), (0.12500000000000000000, 11), (0.0055555555555555555556, 12), (0.000091593624651536418269, 5), (0.56111111111111111111, 9)],
}


def _ratio(num, den):
    # the reference divides the integer literals in Float64 (float(num) / float(den))
    return float(num) / float(den)


def variable_diagonal_sbp_D2(p, N, B, xc=(-1.0, 1.0)):
    """(D, S0, SN, HI, H, M, r) -- diagonal_sbp.jl:470-740 for p = 2, 4, 6:
    D approximates d/dx (B d/dx) with D = HI (-M + SN - S0), M symmetric and
    B-dependent, S0 / SN the B-weighted boundary derivatives."""
    if p not in (2, 4, 6):
        raise ValueError(f"Operators for order {p} are not implemented")
    B = np.asarray(B, dtype=np.float64)
    assert len(B) == N + 1
    Np = N + 1
    b = lambda k: B[k - 1]                         # 1-based access, as the reference
    rows, cols, vals = [], [], []

    def put(i, j, v):                              # 1-based (row, column)
        rows.append(i - 1); cols.append(j - 1); vals.append(v)

    if p == 2:
        bhinv = [2.0]
        BS = np.array([3 / 2, -2.0, 1 / 2])
        put(1, 1, (b(1) + b(2)) / 2)
        put(N + 1, N + 1, (b(N) + b(N + 1)) / 2)
        for i in range(1, N + 1):                  # I_M = 2:N+1, J_M = 1:N
            put(i + 1, i, -(b(i) + b(i + 1)) / 2)
        for i in range(2, N + 1):                  # I_M = J_M = 2:N
            put(i, i, (b(i - 1) + 2 * b(i) + b(i + 1)) / 2)
        for i in range(1, N + 1):                  # I_M = 1:N, J_M = 2:N+1
            put(i, i + 1, -(b(i) + b(i + 1)) / 2)
    elif p == 6:
        bhinv = [43200 / 13649, 8640 / 12013, 4320 / 2711, 4320 / 5359, 8640 / 7877, 43200 / 43801]
        BS = np.array([25 / 12, -4.0, 3.0, -4 / 3, 1 / 4])

        def block6(bk):
            V = np.zeros((9, 9))
            for (r, c), terms in _VAR_P6_M0.items():
                v = 0.0
                for w, k in terms:
                    v += w * bk[k - 1]
                V[r - 1, c - 1] = v
                V[c - 1, r - 1] = v
            return V
        V0 = block6(B[0:12])
        VN = block6(B[::-1][0:12])                 # b1..b12 = B[N+1:-1:N-10]
        for r in range(9):
            for c in range(9):
                put(r + 1, c + 1, V0[r, c])
                put(N + 1 - r, N + 1 - c, VN[r, c])
        for I in range(10, N - 4):                  # J = I-3
            put(I, I - 3, -(11 / 360) * b(I - 3) + (1 / 40) * b(I - 2) + (1 / 40) * b(I - 1) - (11 / 360) * b(I))
        for I in range(10, N - 5):                  # J = I-2
            put(I, I - 2, (1 / 20) * b(I - 3) + (7 / 40) * b(I - 2) - (3 / 10) * b(I - 1) + (7 / 40) * b(I)
                + (1 / 20) * b(I + 1))
        for I in range(10, N - 6):                  # J = I-1
            put(I, I - 1, -(1 / 40) * b(I - 3) - (3 / 10) * b(I - 2) - (17 / 40) * b(I - 1) - (17 / 40) * b(I)
                - (3 / 10) * b(I + 1) - (1 / 40) * b(I + 2))
        for I in range(10, N - 7):                  # J = I
            put(I, I, (1 / 180) * b(I - 3) + (1 / 8) * b(I - 2) + (19 / 20) * b(I - 1) + (101 / 180) * b(I)
                + (19 / 20) * b(I + 1) + (1 / 8) * b(I + 2) + (1 / 180) * b(I + 3))
        for I in range(9, N - 7):                   # J = I+1
            put(I, I + 1, -(1 / 40) * b(I - 2) - (3 / 10) * b(I - 1) - (17 / 40) * b(I) - (17 / 40) * b(I + 1)
                - (3 / 10) * b(I + 2) - (1 / 40) * b(I + 3))
        for I in range(8, N - 7):                   # J = I+2
            put(I, I + 2, (1 / 20) * b(I - 1) + (7 / 40) * b(I) - (3 / 10) * b(I + 1) + (7 / 40) * b(I + 2)
                + (1 / 20) * b(I + 3))
        for I in range(7, N - 7):                   # J = I+3
            put(I, I + 3, -(11 / 360) * b(I) + (1 / 40) * b(I + 1) + (1 / 40) * b(I + 2) - (11 / 360) * b(I + 3))
    else:
        bhinv = [48 / 17, 48 / 59, 48 / 43, 48 / 49]
        BS = np.array([11 / 6, -3.0, 3 / 2, -1 / 3])

        def block(bk):
            V = np.zeros((6, 6))
            for (r, c), terms in _VAR_P4_M0.items():
                v = sum(_ratio(n, d) * bk[k - 1] for n, d, k in terms)
                V[r - 1, c - 1] = v
                V[c - 1, r - 1] = v
            return V
        V0 = block(B[0:8])
        VN = block(B[::-1][0:8])                   # b1..b8 = B[N+1:-1:N-6]
        for r in range(6):
            for c in range(6):
                put(r + 1, c + 1, V0[r, c])
                put(N + 1 - r, N + 1 - c, VN[r, c])   # V_MN[7-r', 7-c'] at rows N-4:N+1
        for J in range(5, N - 4):                   # I = J+2, J = 5:N-5
            put(J + 2, J, (1 / 8) * b(J + 2) - (1 / 6) * b(J + 1) + (1 / 8) * b(J))
        for J in range(6, N - 4):                   # I = J+1, J = 6:N-5
            put(J + 1, J, -(1 / 6) * b(J + 2) - (1 / 2) * b(J + 1) - (1 / 2) * b(J) - (1 / 6) * b(J - 1))
        for J in range(7, N - 4):                   # I = J, J = 7:N-5
            put(J, J, (1 / 24) * b(J + 2) + (5 / 6) * b(J + 1) + (3 / 4) * b(J) + (5 / 6) * b(J - 1)
                + (1 / 24) * b(J - 2))
        for J in range(7, N - 3):                   # I = J-1, J = 7:N-4
            put(J - 1, J, -(1 / 6) * b(J + 1) - (1 / 2) * b(J) - (1 / 2) * b(J - 1) - (1 / 6) * b(J - 2))
        for J in range(7, N - 2):                   # I = J-2, J = 7:N-3
            put(J - 2, J, (1 / 8) * b(J) - (1 / 6) * b(J - 1) + (1 / 8) * b(J - 2))
    bm = len(bhinv)
    if Np < 2 * bm:
        raise ValueError("Grid not big enough to support the operator")
    h = (xc[1] - xc[0]) / N
    assert h > 0
    M = sp.csr_matrix((np.asarray(vals), (rows, cols)), shape=(Np, Np)) / h
    HV = np.ones(Np)
    bh = np.asarray(bhinv)
    HV[:bm] = bh
    HV[Np - bm:] = bh[::-1]
    HI = sp.diags(HV / h).tocsr()
    H = sp.diags(h / HV).tocsr()
    nb = len(BS)
    S0 = sp.csr_matrix((-B[0] * BS / h, (np.zeros(nb, int), np.arange(nb))), shape=(Np, Np))
    SN = sp.csr_matrix((B[-1] * BS / h, (np.full(nb, Np - 1), Np - 1 - np.arange(nb))), shape=(Np, Np))
    D = HI @ (-M + SN - S0)
    r = np.linspace(xc[0], xc[1], Np)
    return D.tocsr(), S0, SN, HI, H, M, r


# ---------------------------------------------------------------------------
# 2-D assembly (level_2_multigrid_new.jl:48-188)
# ---------------------------------------------------------------------------

def _e(i, n):
    v = sp.lil_matrix((n, 1))
    v[i, 0] = 1.0
    return v.tocsr()


def _diag_of(v):
    return sp.diags(np.asarray(v.todense()).ravel()).tocsr()


def operators_2d(level, p=2):
    """Kron-product 2-D operators, level_2_multigrid_new.jl:48-123."""
    Nint = 2 ** level
    D1x, HIx, H1x, _ = diagonal_sbp_D1(p, Nint, xc=(0, 1))
    D2x, S0x, SNx, _, _, _ = diagonal_sbp_D2(p, Nint, xc=(0, 1))
    D1y, HIy, H1y, _ = diagonal_sbp_D1(p, Nint, xc=(0, 1))
    D2y, S0y, SNy, _, _, _ = diagonal_sbp_D2(p, Nint, xc=(0, 1))
    BSx = SNx - S0x
    BSy = SNy - S0y
    n1 = Nint + 1
    e_1x = _e(0, n1); e_Nx = _e(n1 - 1, n1)
    e_1y = _e(0, n1); e_Ny = _e(n1 - 1, n1)
    I = sp.identity(n1, format="csr")
    k = lambda a, b: sp.kron(a, b, format="csr")
    ops = dict(
        e_E=k(e_Nx, I), e_W=k(e_1x, I), e_S=k(I, e_1y), e_N=k(I, e_Ny),
        E_E=k(_diag_of(e_Nx), I), E_W=k(_diag_of(e_1x), I),
        E_S=k(I, _diag_of(e_1y)), E_N=k(I, _diag_of(e_Ny)),
        D1_x=k(D1x, I), D1_y=k(I, D1y),
        D2_x=k(D2x, I), D2_y=k(I, D2y),
        HI_x=k(HIx, I), HI_y=k(I, HIy),
        H_x=k(H1x, I), H_y=k(I, H1y),
        BS_x=k(BSx, I), BS_y=k(I, BSy),
        HI_tilde=k(HIx, HIx), H_tilde=k(H1x, H1y),
    )
    ops["D2"] = ops["D2_x"] + ops["D2_y"]
    return ops


def grid(level):
    N = 2 ** level + 1
    return np.linspace(0.0, 1.0, N)


def sat_operators(level, p=2):
    """SAT matrices and their data maps, level_2_multigrid_new.jl:141-167."""
    o = operators_2d(level, p)
    h = 1.0 / 2 ** level
    tau_W = tau_E = -13.0 / h
    tau_N = tau_S = -1.0
    beta = 1.0
    HI_x, HI_y, BS_x, BS_y = o["HI_x"], o["HI_y"], o["BS_x"], o["BS_y"]
    sat = dict(
        SAT_W=tau_W * HI_x @ o["E_W"] + beta * HI_x @ BS_x.T @ o["E_W"],
        SAT_E=tau_E * HI_x @ o["E_E"] + beta * HI_x @ BS_x.T @ o["E_E"],
        SAT_S=tau_S * HI_y @ o["E_S"] @ BS_y,
        SAT_N=tau_N * HI_y @ o["E_N"] @ BS_y,
        SAT_W_r=tau_W * HI_x @ o["E_W"] @ o["e_W"] + beta * HI_x @ BS_x.T @ o["E_W"] @ o["e_W"],
        SAT_E_r=tau_E * HI_x @ o["E_E"] @ o["e_E"] + beta * HI_x @ BS_x.T @ o["E_E"] @ o["e_E"],
        SAT_S_r=tau_S * HI_y @ o["E_S"] @ o["e_S"],
        SAT_N_r=tau_N * HI_y @ o["E_N"] @ o["e_N"],
    )
    return o, sat


def assemble_operator(level, p=2):
    """SPD A = -H~ (D2 + SAT_W + SAT_E + SAT_S + SAT_N) as CSR
    (level_2_multigrid_new.jl:179,184)."""
    o, s = sat_operators(level, p)
    A = o["D2"] + s["SAT_W"] + s["SAT_E"] + s["SAT_S"] + s["SAT_N"]
    A = -(o["H_tilde"] @ A)
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    return A


def assemble_rhs(level, p, f, gW, gE, gS, gN):
    """b = -H~ (f + SAT_W_r gW + SAT_E_r gE + SAT_S_r gS + SAT_N_r gN)
    (level_2_multigrid_new.jl:181,185).  f is the reference's source, i.e.
    the right side of  Laplace(u) = f ; gW/gE are indexed by the fast index i,
    gS/gN by the slow index j."""
    o, s = sat_operators(level, p)
    b = (np.asarray(f, dtype=np.float64).ravel(order="C")
         + s["SAT_W_r"] @ gW + s["SAT_E_r"] @ gE
         + s["SAT_S_r"] @ gS + s["SAT_N_r"] @ gN)
    return -(o["H_tilde"] @ b)


def rhs_transpose(level, p, lam):
    """Adjoint of ``assemble_rhs`` with respect to (f, gW, gE, gS, gN)."""
    o, s = sat_operators(level, p)
    Hl = o["H_tilde"] @ lam
    return (-Hl, -(s["SAT_W_r"].T @ Hl), -(s["SAT_E_r"].T @ Hl),
            -(s["SAT_S_r"].T @ Hl), -(s["SAT_N_r"].T @ Hl))


def h_tilde_diag(level, p=2):
    o = operators_2d(level, p)
    return o["H_tilde"].diagonal().copy()


# Manufactured solutions --------------------------------------------------

def reference_data(level):
    """u = sin(pi(x+y)) and its data, level_2_multigrid_new.jl:36-38,171-177.
    Returns (f, gW, gE, gS, gN, u) in the flat (i fast) layout."""
    x = grid(level)
    y = x
    U = np.sin(np.pi * x[:, None] + np.pi * y[None, :])   # u(x, y')
    u = U.ravel(order="F")
    f = -2 * np.pi ** 2 * u
    gW = np.sin(np.pi * y)
    gE = -np.sin(np.pi * y)
    gS = -np.pi * np.cos(np.pi * x)
    gN = np.pi * np.cos(np.pi * x + np.pi)
    return f, gW, gE, gS, gN, u


def baseline_data(level):
    """u = sin(pi x) sin(pi y) + x y with x on the slow (Dirichlet) index and
    y on the fast (Neumann) index.  f here is the reference-convention source
    Laplace(u); outward normal derivatives on the Neumann faces."""
    t = grid(level)
    yi = t[:, None]          # fast index -> y
    xj = t[None, :]          # slow index -> x
    U = np.sin(np.pi * xj) * np.sin(np.pi * yi) + xj * yi
    u = U.ravel(order="F")
    f = (-2 * np.pi ** 2 * np.sin(np.pi * xj) * np.sin(np.pi * yi)
         + 0 * yi).ravel(order="F")
    gW = np.zeros_like(t)                       # u(0, y)
    gE = t.copy()                               # u(1, y) = y
    gS = -(np.pi * np.sin(np.pi * t) + t)       # -du/dy at y = 0
    gN = -np.pi * np.sin(np.pi * t) + t         # +du/dy at y = 1
    return f, gW, gE, gS, gN, u


def _var_1d(level, p, b):
    N = 2 ** level
    D, S0, SN, HI, H, M, r = variable_diagonal_sbp_D2(p, N, b, xc=(0, 1))
    return D, (SN - S0).tocsr(), HI, H, M


def assemble_operator_var(level, p, bx, by):
    """SPD A = -H~ (D2x(bx) + D2y(by) + SAT) with one coefficient profile per
    axis: bx along the slow (Dirichlet, x) index, by along the fast
    (Neumann, y) index.  The 2-D construction follows get_operators
    (legacy_functions.jl:100-174) with the variable D2 on both axes, the
    B-weighted boundary derivatives in the SAT terms, and the Dirichlet
    penalty scaled by the face coefficient (tau = -13 b_face / h, the
    reference's -13/h at b = 1)."""
    op, sat = _sat_var(level, p, bx, by)
    A = op["D2"] + sat["SAT_W"] + sat["SAT_E"] + sat["SAT_S"] + sat["SAT_N"]
    A = sp.csr_matrix(-(op["H_tilde"] @ A))
    A.eliminate_zeros()
    A.sort_indices()
    return A


def _sat_var(level, p, bx, by):
    n1 = 2 ** level + 1
    h = 1.0 / 2 ** level
    Dx, BSx1, HIx1, Hx1, _ = _var_1d(level, p, bx)
    Dy, BSy1, HIy1, Hy1, _ = _var_1d(level, p, by)
    I = sp.identity(n1, format="csr")
    k = lambda a, b: sp.kron(a, b, format="csr")
    e1 = _e(0, n1); eN = _e(n1 - 1, n1)
    E_W, E_E = k(_diag_of(e1), I), k(_diag_of(eN), I)
    E_S, E_N = k(I, _diag_of(e1)), k(I, _diag_of(eN))
    HI_x, HI_y = k(HIx1, I), k(I, HIy1)
    BS_x, BS_y = k(BSx1, I), k(I, BSy1)
    tau_W, tau_E = -13.0 * bx[0] / h, -13.0 * bx[-1] / h
    tau_S = tau_N = -1.0
    op = dict(D2=k(Dx, I) + k(I, Dy), H_tilde=k(Hx1, Hy1))
    sat = dict(
        SAT_W=tau_W * HI_x @ E_W + HI_x @ BS_x.T @ E_W,
        SAT_E=tau_E * HI_x @ E_E + HI_x @ BS_x.T @ E_E,
        SAT_S=tau_S * HI_y @ E_S @ BS_y,
        SAT_N=tau_N * HI_y @ E_N @ BS_y,
        SAT_W_r=tau_W * HI_x @ E_W @ k(e1, I) + HI_x @ BS_x.T @ E_W @ k(e1, I),
        SAT_E_r=tau_E * HI_x @ E_E @ k(eN, I) + HI_x @ BS_x.T @ E_E @ k(eN, I),
        SAT_S_r=tau_S * HI_y @ E_S @ k(I, e1),
        SAT_N_r=tau_N * HI_y @ E_N @ k(I, eN),
    )
    return op, sat


def assemble_rhs_var(level, p, bx, by, f, gW, gE, gS, gN):
    """b = -H~ (f + SAT_W_r gW + SAT_E_r gE + SAT_S_r gS + SAT_N_r gN) of the
    variable-coefficient operator; gS / gN are the outward fluxes by du/dn."""
    op, s = _sat_var(level, p, bx, by)
    b = (np.asarray(f, dtype=np.float64).ravel(order="C") + s["SAT_W_r"] @ gW + s["SAT_E_r"] @ gE
         + s["SAT_S_r"] @ gS + s["SAT_N_r"] @ gN)
    return -(op["H_tilde"] @ b)


def var_coefficients(level):
    """Smooth positive profiles for tests: bx(x) = 1 + x / 2 + sin(3x) / 4 on the
    slow index, by(y) = 1 + y^2 on the fast index."""
    t = grid(level)
    return 1.0 + t / 2 + np.sin(3 * t) / 4, 1.0 + t ** 2


def var_data(level):
    """u = sin(pi x) sin(pi y) + x y with the profiles of var_coefficients:
    f = d/dx (bx u_x) + d/dy (by u_y), Dirichlet gW / gE = u, Neumann fluxes
    gS = -by u_y at y = 0, gN = by u_y at y = 1.  Returns (bx, by, f, gW, gE, gS, gN, u)."""
    t = grid(level)
    yi = t[:, None]
    xj = t[None, :]
    bx = 1.0 + xj / 2 + np.sin(3 * xj) / 4
    dbx = 0.5 + 0.75 * np.cos(3 * xj)
    by = 1.0 + yi ** 2
    dby = 2 * yi
    pi = np.pi
    U = np.sin(pi * xj) * np.sin(pi * yi) + xj * yi
    ux = pi * np.cos(pi * xj) * np.sin(pi * yi) + yi
    uxx = -pi ** 2 * np.sin(pi * xj) * np.sin(pi * yi)
    uy = pi * np.sin(pi * xj) * np.cos(pi * yi) + xj
    uyy = -pi ** 2 * np.sin(pi * xj) * np.sin(pi * yi)
    F = dbx * ux + bx * uxx + dby * uy + by * uyy
    bxs, bys = var_coefficients(level)
    gW = t * 0.0                                   # u(0, y) = 0
    gE = t.copy()                                  # u(1, y) = y
    gS = -(1.0 + 0.0) * (pi * np.sin(pi * t) + t)  # -by(0) u_y(x, 0)
    gN = (1.0 + 1.0) * (-pi * np.sin(pi * t) + t)  #  by(1) u_y(x, 1)
    return bxs, bys, F.ravel(order="F"), gW, gE, gS, gN, U.ravel(order="F")


def one_d_factors_var(p, N, b):
    """Dimensionless variable-coefficient factors (hw, m, k): m = h M(b),
    k = m - (E bs_b + bs_b^T E) + 13 b_face E with bs_b the b-weighted
    boundary rows, so that A = kron(k(bx), diag(hw)) + kron(diag(hw), m(by))."""
    D, S0, SN, HI, H, M, r = variable_diagonal_sbp_D2(p, N - 1, b, xc=(0, 1))
    h = 1.0 / (N - 1)
    m = h * M.toarray()
    BS = h * (SN - S0).toarray()
    E = np.zeros((N, N)); E[0, 0] = b[0]; E[-1, -1] = b[-1]
    k = m - (np.diag((np.arange(N) == 0) | (np.arange(N) == N - 1)).astype(float) @ BS + BS.T @ np.diag(
        ((np.arange(N) == 0) | (np.arange(N) == N - 1)).astype(float))) + 13.0 * E
    return H.diagonal() / h, m, k


def assembling_matrix(level, p=2):
    """(A, b, H_tilde, Nx, Ny, analy_sol) -- level_2_multigrid_new.jl:125-188."""
    A = assemble_operator(level, p)
    f, gW, gE, gS, gN, u = reference_data(level)
    b = assemble_rhs(level, p, f, gW, gE, gS, gN)
    o = operators_2d(level, p)
    N = 2 ** level + 1
    return A, b, o["H_tilde"], N, N, u


# ---------------------------------------------------------------------------
# Literal restatement of the reference GPU apply (p = 2), returns -A u
# ---------------------------------------------------------------------------

def reference_gpu_apply_p2(u_flat, N):
    """matrix_free_A_full_GPU (split_matrix_free.jl:444-486) written out with
    numpy: D2_split_naive_v2 (:32-49) + SBP_N!/SBP_S!/SBP_W!/SBP_E!
    (:324-429) + the four edge fix-ups (:478-483).  Arrays are indexed
    [i, j] with i fast, as in the Julia column-major (Nx, Ny) arrays."""
    U = np.asarray(u_flat, dtype=np.float64).reshape(N, N, order="F")
    Nx = Ny = N
    h = 1.0 / (Nx - 1)
    a1 = a2 = -13.0 / h
    a3 = a4 = -1.0
    beta = 1.0
    out = np.zeros((Nx, Ny))
    # D2_split_naive_v2: interior 5-point, zero on the boundary
    out[1:-1, 1:-1] = (U[:-2, 1:-1] + U[2:, 1:-1] + U[1:-1, :-2] + U[1:-1, 2:]
                       - 4 * U[1:-1, 1:-1])

    def sbp_n(I):   # I = idata[1:3, :] (3 x Ny) -> 1 x Ny
        o = np.zeros(Ny)
        for idx in range(1, Ny + 1):
            c = idx - 1
            if 4 <= idx <= Ny - 3:
                o[c] = (I[0, c] - 2 * I[1, c] + I[2, c] + I[0, c - 1] - 2 * I[0, c]
                        + I[0, c + 1] + 2 * a3 * (1.5 * I[0, c] - 2 * I[1, c] + 0.5 * I[2, c])) / 2
            if idx == 1:
                o[c] = (I[0, c] - 2 * I[0, c + 1] + I[0, c + 2] + I[0, c] - 2 * I[1, c] + I[2, c]
                        + 2 * a3 * (1.5 * I[0, c] - 2 * I[1, c] + 0.5 * I[2, c])) / 4
                for q in (c + 1, c + 2):
                    o[q] = (I[0, q] - 2 * I[1, q] + I[2, q] + I[0, q - 1] - 2 * I[0, q]
                            + I[0, q + 1] + 2 * a3 * (1.5 * I[0, q] - 2 * I[1, q] + 0.5 * I[2, q])) / 2
            if idx == Ny:
                o[c] = (I[0, c] - 2 * I[0, c - 1] + I[0, c - 2] + I[0, c] - 2 * I[1, c] + I[2, c]
                        + 2 * a3 * (1.5 * I[0, c] - 2 * I[1, c] + 0.5 * I[2, c])) / 4
                for q in (c - 1, c - 2):
                    o[q] = (I[0, q] - 2 * I[1, q] + I[2, q] + I[0, q - 1] - 2 * I[0, q]
                            + I[0, q + 1] + 2 * a3 * (1.5 * I[0, q] - 2 * I[1, q] + 0.5 * I[2, q])) / 2
        return o

    def sbp_s(I):   # I = idata[Nx-2:Nx, :]; row 3 is the boundary row
        o = np.zeros(Ny)
        for idx in range(1, Ny + 1):
            c = idx - 1
            if 4 <= idx <= Ny - 3:
                o[c] = (I[2, c] - 2 * I[1, c] + I[0, c] + I[2, c - 1] - 2 * I[2, c]
                        + I[2, c + 1] + 2 * a4 * (1.5 * I[2, c] - 2 * I[1, c] + 0.5 * I[0, c])) / 2
            if idx == 1:
                o[c] = (I[2, c] - 2 * I[2, c + 1] + I[2, c + 2] + I[2, c] - 2 * I[1, c] + I[0, c]
                        + 2 * a4 * (1.5 * I[2, c] - 2 * I[1, c] + 0.5 * I[0, c])) / 4
                for q in (c + 1, c + 2):
                    o[q] = (I[2, q] - 2 * I[1, q] + I[0, q] + I[2, q - 1] - 2 * I[2, q]
                            + I[2, q + 1] + 2 * a4 * (1.5 * I[2, q] - 2 * I[1, q] + 0.5 * I[0, q])) / 2
            if idx == Ny:
                o[c] = (I[2, c] - 2 * I[2, c - 1] + I[2, c - 2] + I[2, c] - 2 * I[1, c] + I[0, c]
                        + 2 * a4 * (1.5 * I[2, c] - 2 * I[1, c] + 0.5 * I[0, c])) / 4
                for q in (c - 1, c - 2):
                    o[q] = (I[2, q] - 2 * I[1, q] + I[0, q] + I[2, q - 1] - 2 * I[2, q]
                            + I[2, q + 1] + 2 * a4 * (1.5 * I[2, q] - 2 * I[1, q] + 0.5 * I[0, q])) / 2
        return o

    def sbp_w(I):   # I = idata[:, 1:3] (Nx x 3) -> Nx x 3
        o = np.zeros((Nx, 3))
        for idx in range(1, Nx + 1):
            r = idx - 1
            if 2 <= idx <= Nx - 1:
                o[r, 0] = (I[r - 1, 0] - 2 * I[r, 0] + I[r + 1, 0] + I[r, 0] - 2 * I[r, 1] + I[r, 2]
                           + 2 * beta * (1.5 * I[r, 0]) + 2 * a1 * I[r, 0] * h) / 2
                o[r, 1] = 2 * beta * (-1 * I[r, 0])
                o[r, 2] = 0.5 * beta * I[r, 0]
            if idx == 1 or idx == Nx:
                o[r, 0] = (2 * beta * (1.5 * I[r, 0]) + 2 * a1 * I[r, 0] * h) / 4
                o[r, 1] = 2 * beta * (-1 * I[r, 0]) / 2
                o[r, 2] = 0.5 * beta * I[r, 0] / 2
        return o

    def sbp_e(I):   # I = idata[:, Ny-2:Ny]; column 3 is the boundary column
        o = np.zeros((Nx, 3))
        for idx in range(1, Nx + 1):
            r = idx - 1
            if 2 <= idx <= Nx - 1:
                o[r, 2] = (I[r - 1, 2] - 2 * I[r, 2] + I[r + 1, 2] + I[r, 2] - 2 * I[r, 1] + I[r, 0]
                           + 2 * beta * (1.5 * I[r, 2]) + 2 * a2 * I[r, 2] * h) / 2
                o[r, 1] = 2 * beta * (-1 * I[r, 2])
                o[r, 0] = 0.5 * beta * I[r, 2]
            if idx == 1 or idx == Nx:
                o[r, 2] = (2 * beta * (1.5 * I[r, 2]) + 2 * a2 * I[r, 2] * h) / 4
                o[r, 1] = 2 * beta * (-1 * I[r, 2]) / 2
                o[r, 0] = 0.5 * beta * I[r, 2] / 2
        return o

    on = sbp_n(U[0:3, :])
    os_ = sbp_s(U[Nx - 3:Nx, :])
    ow = sbp_w(U[:, 0:3])
    oe = sbp_e(U[:, Ny - 3:Ny])
    out[:, 0:3] += ow
    out[:, Ny - 3:Ny] += oe
    out[0, :] += on
    out[Nx - 1, :] += os_
    return out.ravel(order="F")


# ---------------------------------------------------------------------------
# Closed form A = K (x) H + H (x) M  (dimensionless 1-D factors)
# ---------------------------------------------------------------------------

def one_d_factors(p, N):
    """Dimensionless 1-D factors (hw, m, k) with
    A[(i,j),(i',j')] = hw[i] k[j,j'] d(i,i') + hw[j] m[i,i'] d(j,j').
    m = h (BS - H D2) and k = m - (E bs + bs^T E) + 13 E, built here from the
    D2 matrices of ``diagonal_sbp_D2``."""
    D2, S0, SN, HI, H, _ = diagonal_sbp_D2(p, N - 1, xc=(0, 1))
    h = 1.0 / (N - 1)
    BS = (SN - S0).toarray()
    Hd = H.diagonal()
    m = h * (BS - (sp.diags(Hd) @ D2).toarray())
    E = np.zeros((N, N)); E[0, 0] = 1; E[-1, -1] = 1
    k = m - (E @ (h * BS) + (h * BS).T @ E) + 13.0 * E
    return Hd / h, m, k


def closed_form_operator(level, p):
    N = 2 ** level + 1
    hw, m, k = one_d_factors(p, N)
    # kron(k, diag(hw)) acts on the slow index with k, weighted by hw[i]
    A = sp.kron(sp.csr_matrix(k), sp.diags(hw)) + sp.kron(sp.diags(hw), sp.csr_matrix(m))
    return sp.csr_matrix(A)


# ---------------------------------------------------------------------------
# Transfer operators (level_2_multigrid_new.jl:190-246)
# ---------------------------------------------------------------------------

def prolongation_matrix(N):
    P = sp.lil_matrix((2 * N - 1, N))
    for i in range(1, 2 * N):
        if i % 2 == 1:
            P[i - 1, (i + 1) // 2 - 1] = 1.0
        else:
            P[i - 1, i // 2 - 1] = 0.5
            P[i - 1, i // 2] = 0.5
    return P.tocsr()


def restriction_matrix_normal(N):
    nc = (N + 1) // 2
    R = sp.lil_matrix((nc, N))
    R[0, 0] = 0.5; R[0, 1] = 0.25
    R[nc - 1, N - 2] = 0.25; R[nc - 1, N - 1] = 0.5
    for i in range(2, nc):
        R[i - 1, 2 * i - 3] = 0.25
        R[i - 1, 2 * i - 2] = 0.5
        R[i - 1, 2 * i - 1] = 0.25
    return R.tocsr()


def prolongation_2d(Nc):
    P1 = prolongation_matrix(Nc)
    return sp.kron(P1, P1, format="csr")


def restriction_2d(Nf):
    R1 = restriction_matrix_normal(Nf)
    return sp.kron(R1, R1, format="csr")


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------

def cg_cpu(A, b, x=None, maxiter=None, abstol=None):
    """CG_CPU, level_2_multigrid_new.jl:258-291: stops when ||r||_2 < abstol.
    Returns (x, iterations, norms)."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x is None else np.array(x, dtype=np.float64)
    if maxiter is None:
        maxiter = len(b)
    if abstol is None:
        abstol = math.sqrt(np.finfo(np.float64).eps)
    r = b - A @ x
    p = r.copy()
    rsold = r @ r
    norms = [math.sqrt(rsold)]
    it = 0
    for _ in range(maxiter):
        it += 1
        Ap = A @ p
        alpha = rsold / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        rsnew = r @ r
        norms.append(math.sqrt(rsnew))
        if math.sqrt(rsnew) < abstol:
            break
        p = r + (rsnew / rsold) * p
        rsold = rsnew
    return x, it, norms


def jacobi_sweeps(A, x, b, nu, omega=2.0 / 3.0, dinv=None):
    """jacobi_brittany!, level_2_multigrid_new.jl:335-344:
    x <- omega D^-1 (b - Q x) + (1 - omega) x  ==  x + omega D^-1 (b - A x)."""
    if dinv is None:
        dinv = 1.0 / A.diagonal()
    for _ in range(nu):
        x = x + omega * dinv * (b - A @ x)
    return x


def modified_richardson(A, x, b, nu=3, omega=0.15):
    """level_2_multigrid_new.jl:346-350."""
    for _ in range(nu):
        x = x + omega * (b - A @ x)
    return x


class Hierarchy:
    """Rediscretised operators A_l (level_2_multigrid_new.jl:125-188 at each
    level, as initial_guess_cg.jl:514-521 does for the coarse operator) from
    the fine level down to 17x17, with transfers P_l and R_l."""

    def __init__(self, level, p, coarse_level=4, rscale=4.0):
        self.levels = list(range(level, coarse_level - 1, -1))
        self.A = [assemble_operator(l, p) for l in self.levels]
        self.dinv = [1.0 / A.diagonal() for A in self.A]
        self.P = [prolongation_2d(2 ** l + 1) for l in self.levels[1:]]
        # rscale * full weighting; rscale = 4 gives R = P^T
        self.R = [rscale * restriction_2d(2 ** l + 1) for l in self.levels[:-1]]
        self.coarse = spla.factorized(sp.csc_matrix(self.A[-1]))


def vcycle(hier, r, nu_pre=2, nu_post=2, omega=2.0 / 3.0, lvl=0):
    """One V-cycle z ~ A^-1 r from a zero initial guess: nu_pre damped-Jacobi
    sweeps, residual, restriction, recursive coarse correction, bilinear
    prolongation, nu_post sweeps; exact solve on the coarsest level."""
    if lvl == len(hier.A) - 1:
        return hier.coarse(r)
    A = hier.A[lvl]
    x = np.zeros_like(r)
    x = jacobi_sweeps(A, x, r, nu_pre, omega, hier.dinv[lvl])
    rc = hier.R[lvl] @ (r - A @ x)
    ec = vcycle(hier, rc, nu_pre, nu_post, omega, lvl + 1)
    x = x + hier.P[lvl] @ ec
    x = jacobi_sweeps(A, x, r, nu_post, omega, hier.dinv[lvl])
    return x


def mgpcg(hier, b, x=None, rtol=1e-8, maxiter=200, nu_pre=2, nu_post=2, omega=2.0 / 3.0):
    """mg_preconditioned_CG (level_2_multigrid_new.jl:480-524) with the
    multi-level V-cycle as preconditioner; stops when ||r||_2 < rtol ||b||_2.
    Returns (x, iterations, norms)."""
    A = hier.A[0]
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x is None else np.array(x, dtype=np.float64)
    abstol = rtol * np.linalg.norm(b)
    r = b - A @ x
    z = vcycle(hier, r, nu_pre, nu_post, omega)
    p = z.copy()
    rz = r @ z
    norms = [np.linalg.norm(r)]
    it = 0
    for _ in range(maxiter):
        it += 1
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        rs = r @ r
        norms.append(math.sqrt(rs))
        if math.sqrt(rs) < abstol:
            break
        z = vcycle(hier, r, nu_pre, nu_post, omega)
        rznew = r @ z
        p = z + (rznew / rz) * p
        rz = rznew
    return x, it, norms


def initial_guess_interpolation_cg(A, b, A2h, b2h, abstol, maxiter=None, A4h=None, b4h=None):
    """initial_guess_interpolation_CG (initial_guess_cg.jl:394-399): x_2h =
    A_2h \\ b_2h (sparse direct solve), x0 = P x_2h (bilinear), then CG on
    A x = b from x0 to ||r|| < abstol.  With A4h/b4h, the three-level form
    (initial_guess_cg.jl:447-456): the direct solve is on the 4h grid and CG
    on the 2h grid from P x_4h (same abstol) supplies x_2h.
    Returns (x, iterations per CG grid (coarse to fine), final ||r|| per CG
    grid, x0)."""
    N = int(round(math.sqrt(len(b))))
    N2 = (N + 1) // 2
    iters, norms = [], []
    if A4h is None:
        x2 = spla.spsolve(sp.csc_matrix(A2h), np.asarray(b2h, dtype=np.float64))
    else:
        N4 = (N2 + 1) // 2
        x4 = spla.spsolve(sp.csc_matrix(A4h), np.asarray(b4h, dtype=np.float64))
        x2, it2, nr2 = cg_cpu(A2h, b2h, prolongation_2d(N4) @ x4, maxiter=maxiter, abstol=abstol)
        iters.append(it2)
        norms.append(nr2[-1])
    x0 = prolongation_2d(N2) @ x2
    x, it, nr = cg_cpu(A, b, x0.copy(), maxiter=maxiter, abstol=abstol)
    iters.append(it)
    norms.append(nr[-1])
    return x, iters, norms, x0


def bilinear_data(level, c=(0.5, -1.0, 2.0, 3.0)):
    """u = c0 + c1 x + c2 y + c3 x y (x slow / Dirichlet, y fast / Neumann)
    and its data: f = Laplace(u) = 0, gW/gE = u on x = 0/1, gS = -du/dy at
    y = 0, gN = du/dy at y = 1.  SBP D2 and the boundary derivative are exact
    on linear functions, so u is the discrete solution on every grid."""
    t = grid(level)
    yi = t[:, None]
    xj = t[None, :]
    U = c[0] + c[1] * xj + c[2] * yi + c[3] * xj * yi
    u = U.ravel(order="F")
    f = np.zeros_like(u)
    gW = c[0] + c[2] * t
    gE = c[0] + c[1] + (c[2] + c[3]) * t
    gS = -(c[2] + c[3] * t)
    gN = c[2] + c[3] * t
    return f, gW, gE, gS, gN, u


def two_level_richardson(A, A2h, b, nu=3, omega=0.15):
    """matrix_free_Two_level_multigrid (initial_guess_cg.jl:147-180) on the
    SPD A (the reference's -A kernel and sign flips cancel): nu Richardson
    sweeps from zero, residual, full weighting (no scaling), exact coarse
    solve with the rediscretised A_2h, bilinear prolongation, nu sweeps.
    Returns (z, ||A z - b||)."""
    N = int(round(math.sqrt(len(b))))
    N2 = (N + 1) // 2
    v = modified_richardson(A, np.zeros_like(b), b, nu, omega)
    f = restriction_2d(N) @ (b - A @ v)
    v = v + prolongation_2d(N2) @ spla.spsolve(sp.csc_matrix(A2h), f)
    v = modified_richardson(A, v, b, nu, omega)
    return v, float(np.linalg.norm(A @ v - b))


def matrix_free_mgcg(A, b, precond, x=None, abstol=None, maxiter=None):
    """matrix_free_MGCG (initial_guess_cg.jl:303-346) with a given
    preconditioner z = precond(r); stops when ||r||_2 < abstol.
    Returns (x, iterations, norms)."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x is None else np.array(x, dtype=np.float64)
    abstol = math.sqrt(np.finfo(np.float64).eps) if abstol is None else abstol
    maxiter = len(b) if maxiter is None else maxiter
    r = b - A @ x
    z = precond(r)
    p = z.copy()
    rz = r @ z
    norms = [float(np.linalg.norm(r))]
    it = 0
    for _ in range(maxiter):
        it += 1
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        rs = r @ r
        norms.append(math.sqrt(rs))
        if math.sqrt(rs) < abstol:
            break
        z = precond(r)
        rznew = r @ z
        p = z + (rznew / rz) * p
        rz = rznew
    return x, it, norms


def h_norm(v, hdiag):
    """sqrt(v' H~ v) (matrix_free_mg_cg.jl:464,490)."""
    return math.sqrt(float(v @ (hdiag * v)))


# ---------------------------------------------------------------------------
# Seeded Fourier-mode data (random source and boundary data)
# ---------------------------------------------------------------------------

def splitmix64(state):
    """One splitmix64 step on a Python int; returns (new_state, value)."""
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def fourier_modes(seed, n_src=64, n_bnd=8, kmax=32):
    """Mode lists for the random source f and the four boundary data vectors.
    f = sum_m a_m cos(pi (kx x + ky y) + phi_m), a_m = 1/(kx^2+ky^2+1);
    each g = sum_m a_m cos(pi k t + phi_m), a_m = 1/(k^2+1).
    Integers kx, ky, k in 1..kmax and phases phi in [0, 2pi) are drawn with
    splitmix64 from ``seed`` in a fixed order (source first, then W, E, S, N)."""
    st = seed & 0xFFFFFFFFFFFFFFFF
    def u01():
        nonlocal st
        st, v = splitmix64(st)
        return (v >> 11) * (1.0 / 9007199254740992.0)
    def kint():
        nonlocal st
        st, v = splitmix64(st)
        return 1 + (v % kmax)
    src = []
    for _ in range(n_src):
        kx = kint(); ky = kint(); ph = 2 * np.pi * u01()
        src.append((kx, ky, 1.0 / (kx * kx + ky * ky + 1.0), ph))
    bnd = []
    for _ in range(4):
        modes = []
        for _ in range(n_bnd):
            kk = kint(); ph = 2 * np.pi * u01()
            modes.append((kk, 1.0 / (kk * kk + 1.0), ph))
        bnd.append(modes)
    return src, bnd


def fourier_data(level, seed):
    """(f, gW, gE, gS, gN) for the seeded Fourier-mode problem on level."""
    t = grid(level)
    src, bnd = fourier_modes(seed)
    yi = t[:, None]
    xj = t[None, :]
    F = np.zeros((len(t), len(t)))
    for kx, ky, a, ph in src:
        F += a * np.cos(np.pi * (kx * xj + ky * yi) + ph)
    g = []
    for modes in bnd:
        v = np.zeros_like(t)
        for kk, a, ph in modes:
            v += a * np.cos(np.pi * kk * t + ph)
        g.append(v)
    return (F.ravel(order="F"), g[0], g[1], g[2], g[3])

"""A numpy statement of HBP_ISP::run_ms (hbp_isp.cu:1313-1346), written from the reference's
text to check the oracle's BP independently of it.  Test infrastructure only: tiny pyramids,
plain loops where the reference has per-thread loops, float32 arithmetic in the reference's
order of operations (numpy rounds every float32 operation and never contracts).

Where the reference's result depends on its atomics' order, on uninitialised memory or on a
division by zero, the statement follows the rules the library pins (its own deviations, named
below where they are used):
  * plane-cost cluster sums: each pixel's lambda*cost is rounded to an integer multiple of
    2^-shift and the multiples are summed exactly (any order gives the same sum);
  * belief sums: cost, then the messages in CSR (edge) order; dc0 sums |dL|, |da|, |db| in
    that order; dist_pos sums dx^2 then dy^2;
  * the top level's planes (never written by the reference) are zero;
  * map_msg's atomicMinFloat is a left fold over the slots with NaN read as +qNaN.

Every function cites the reference lines it states.
"""
import math

import numpy as np

F = np.float32
P = 4
K = 3 * P
FLT_MAX = F(np.finfo(np.float32).max)
FLT_EPSILON = F(np.finfo(np.float32).eps)
GAMMA = 20.0           # m_gamma (hbp_isp.hpp:278), a double
END_DZ = F(0.01)       # end_dz (hbp_isp.hpp:265)
INLIER = F(1.0)        # inlier_thrsh (hbp_isp.hpp:277)
M31 = 1 << 31


def lcg(s):
    """rand_lcg (hbp_isp.cu:551-559): (1103515245 s + 12345) mod 2^31, in uint32 arithmetic."""
    return ((1103515245 * s + 12345) & 0xFFFFFFFF) % M31


def lcg_uniform(s, f):
    """rand_lcg_uniform (hbp_isp.cu:560-569): the next state, and s / (float)2^31 * f in float."""
    s = lcg(s)
    return s, F(F(F(s) / F(M31)) * F(f))


def umod(s, n):
    """s % n with s unsigned and n an int converted to unsigned (n = 0: the pinned 0)."""
    n &= 0xFFFFFFFF
    return 0 if n == 0 else s % n


def three_diff(p0, p1, p2, tn):
    """make_three_diff_id (hbp_isp.cu:504-529)."""
    if tn <= 3:
        p1 = (p0 + 1) % tn
        p2 = (p1 + 1) % tn
        return p0, p1, p2
    if p1 == p0:
        p1 = tn - 1
    if p2 == p0:
        p2 = tn - 1 if p1 == tn - 2 else tn - 2
    elif p2 == p1:
        p2 = tn - 1 if p0 == tn - 2 else tn - 2
    return p0, p1, p2


def fit_plane(A, B, C):
    """fit_plane_from_points (hbp_isp.cu:531-548) on float32 4-vectors."""
    with np.errstate(all="ignore"):
        a = [F(B[i] - A[i]) for i in range(4)]
        b = [F(C[i] - A[i]) for i in range(4)]
        nx = F(F(a[1] * b[2]) - F(a[2] * b[1]))
        ny = F(F(a[2] * b[0]) - F(a[0] * b[2]))
        nz = F(F(a[0] * b[1]) - F(a[1] * b[0]))
        nx = F(nx / nz)
        ny = F(ny / nz)
        nz = F(nz / nz)
        if abs(F(nz - F(1.0))) > FLT_EPSILON:
            return np.array([0, 0, F(F(F(A[2] + B[2]) + C[2]) / F(3.0)), 0], np.float32)
        return np.array([-nx, -ny, F(F(F(nx * A[0]) + F(ny * A[1])) + A[2]), 0], np.float32)


def floor_to_int(df):
    """__float2int_rd (hbp_isp.cu:217): NaN -> 0, saturating at the int range (as an int64)."""
    with np.errstate(invalid="ignore"):
        fl = np.floor(df.astype(np.float64))
        fl = np.where(np.isnan(fl), 0.0, fl)
        fl = np.clip(fl, -2.0 ** 31, 2.0 ** 31 - 1)
    return fl.astype(np.int64)


def fixed_shift(lam, tau, use_cost_vol, npix):
    """The plane-cost sums' fixed-point scale: the largest term's binary exponent plus the pixel
    count's stay below 52 bits."""
    vmax = abs(float(lam)) * (abs(float(tau)) if use_cost_vol else 1.0) * 1.0001
    ev = math.frexp(vmax)[1] if vmax > 0 else 0
    en = math.frexp(float(max(npix, 1)))[1]
    return int(min(120, max(-60, 52 - ev - en)))


def amin_fold(stored, v):
    """atomicMinFloat (util.h:141-148) as one step of a left fold; NaN read as +qNaN."""
    sb = np.float32(stored).view(np.uint32).item()
    vb = 0x7FC00000 if v != v else np.float32(v).view(np.uint32).item()
    if v >= 0:
        si = sb - (1 << 32) if sb >= 1 << 31 else sb
        vi = vb - (1 << 32) if vb >= 1 << 31 else vb
        r = min(si, vi) & 0xFFFFFFFF
    else:
        r = max(sb, vb)
    return np.array([r], np.uint32).view(np.float32)[0]


class Level:
    """One pyramid level with the prologue of run_ms (hbp_isp.cu:1319-1325)."""

    def __init__(self, lv, im_clus, W, N):
        self.V, self.E = lv["V"], lv["E"]
        V, E = self.V, self.E
        self.clus = lv["clus"].astype(np.int64)
        self.bd = lv["bd"].astype(np.int64).reshape(E, 2)
        self.bdl = lv["bdl"].astype(np.int64)
        # sum_to_mean (segment.cu:492-503): the first five channels divided by the weight
        m = lv["mean"].astype(np.float32).reshape(6, V).copy()
        with np.errstate(all="ignore"):
            for ch in range(5):
                m[ch] = m[ch] / m[5]
        self.mean = m
        # reverse_bd_id (hbp_isp.cu:1003-1019, 140-175): CSR offsets by source, reverse edges
        cnt = np.bincount(self.bd[:, 0], minlength=V) if E else np.zeros(V, np.int64)
        self.nb_num = cnt.astype(np.int64)
        self.nb_pos = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        self.rev = np.full(E, -1, np.int64)
        for e in range(E):
            s, t = self.bd[e]
            for r in range(self.nb_pos[t], self.nb_pos[t + 1]):
                if self.bd[r, 1] == s:
                    self.rev[e] = r
                    break
        # compute_dist_color_pos (hbp_isp.cu:952-970, 78-126)
        s, t = self.bd[:, 0], self.bd[:, 1]
        dcol = np.zeros(E, np.float32)
        for ch in range(3):
            dcol = (dcol + np.abs(m[ch][s] - m[ch][t])).astype(np.float32)
        self.dc0 = np.array([F(math.exp(-float(d) / GAMMA)) for d in dcol], np.float32)
        dpos = np.zeros(E, np.float32)
        for ch in (3, 4):
            dd = m[ch][s] - m[ch][t]
            dpos = dpos + dd * dd
        self.dc1 = np.sqrt(dpos).astype(np.float32)
        # pixels sorted by cluster, ascending pixel id within (setup_D4, hbp_isp.cu:871-891)
        self.px = np.argsort(im_clus, kind="stable").astype(np.int64)
        self.px_clus = im_clus[self.px]
        pc = np.bincount(im_clus, minlength=V)
        self.pix_num = pc.astype(np.int64)
        self.pix_pos = np.concatenate([[0], np.cumsum(pc)[:-1]]).astype(np.int64)
        # sort_clus (hbp_isp.cu:985-1001): nodes grouped by parent, ascending within
        self.clus_sort = np.argsort(self.clus, kind="stable").astype(np.int64)
        # per-level state
        self.plane = np.zeros((K, V, 4), np.float32)
        self.cost = np.full((K, V), np.nan, np.float32)
        self.belief = np.full((K, V), np.nan, np.float32)
        self.point = np.zeros((V, 3), np.int64)


class HBP:
    """HBP_ISP (hbp_isp.hpp:288-386) with run_ms split into sweeps."""

    def __init__(self, levels, vol, dwta, W, H, D, alpha=1000.1, tau_s=0.6, min_level=2, use_cost_vol=1,
                 tau=25.0):
        self.W, self.H, self.D = W, H, D
        self.N = W * H
        self.L = len(levels)
        self.lam, self.tau_s, self.tau = F(alpha), F(tau_s), F(tau)
        self.use_cost_vol = use_cost_vol
        self.min_level = min_level
        self.vol = vol.reshape(D, self.N).astype(np.float32)
        self.dwta = dwta.reshape(self.N).astype(np.float32)
        self.shift = fixed_shift(self.lam, self.tau, use_cost_vol, self.N)
        im = np.arange(self.N, dtype=np.int64)
        self.lv = []
        for l, lv in enumerate(levels):
            self.lv.append(Level(lv, im, W, self.N))
            im = lv["clus"].astype(np.int64)[im]
        # the top level: planes zero, messages and beliefs zero (hbp_isp.cu:1334-1338)
        top = self.lv[-1]
        top.belief[:] = 0
        # seeds shared by all levels (hbp_isp.hpp:368-370), written per level by init_bp
        n0 = self.lv[0].V
        self.s_child = np.zeros(n0, np.int64)
        self.s_neighbor = np.zeros(n0, np.int64)
        self.s_ransac = np.zeros(n0, np.int64)
        self.msg = None
        self.sched = self._schedule()
        self.next = 0

    # ---------------------------------------------------------------- schedule
    def _schedule(self):
        """hbp_g (hbp_isp.cu:1283-1311) for levels L-2 .. min_level: (level, kind, dz, dn, first)."""
        out = []
        for l in range(self.L - 2, self.min_level - 1, -1):
            iter_bp = 10 + l * 10
            max_dz = F(F(F(self.D) / F(2.0)) * F(l)) / F(self.L)
            max_dz = F(max_dz)
            max_dn = F(F(F(1.0) * F(l)) / F(self.L))
            first = True
            for i in range(iter_bp):
                rand_ok = not (i <= iter_bp // 4 * 3)
                for it in range(6):
                    if it <= 3:
                        kind = ("ransac", None, None)
                    elif it <= 4:
                        kind = ("spatial", None, None)
                    elif it <= 5 and rand_ok and max_dz >= END_DZ:
                        kind = ("random", max_dz, max_dn)
                        max_dz = F(max_dz / F(2.0))
                        max_dn = F(max_dn / F(2.0))
                    else:
                        break
                    out.append((l, kind[0], kind[1], kind[2], first))
                    first = False
        return out

    # ---------------------------------------------------------------- plane cost
    def plane_cost(self, l, s_c):
        """compute_plane_cost (hbp_isp.cu:1159-1174) with read_plane_cost (177-228), the cluster
        sums in the pinned fixed point."""
        B = self.lv[l]
        W = self.W
        x = (B.px % W).astype(np.float32)
        y = (B.px // W).astype(np.float32)
        c = B.px_clus
        for p in range(P):
            pl = B.plane[s_c + p][c]
            with np.errstate(all="ignore"):
                df = (pl[:, 0] * x + pl[:, 1] * y) + pl[:, 2]
                if self.use_cost_vol:
                    d = floor_to_int(df)
                    out = (d < 0) | (d + 1 >= self.D)
                    dc = np.clip(d, 0, self.D - 2)
                    c0 = self.vol[dc, B.px]
                    c1 = self.vol[dc + 1, B.px]
                    v = (d + 1).astype(np.float32) - df
                    w = df - d.astype(np.float32)
                    cost = v * c0 + w * c1
                    cost = np.where(out, self.tau, cost).astype(np.float32)
                else:
                    cost = np.where(np.abs(self.dwta[B.px] - df) <= INLIER, F(0), F(1)).astype(np.float32)
                val = (self.lam * cost).astype(np.float32)
            nan = np.isnan(val)
            q = np.rint(np.ldexp(np.where(nan, 0.0, val.astype(np.float64)), self.shift)).astype(np.int64)
            sums = np.zeros(B.V, np.int64)
            np.add.at(sums, c, q)
            has_nan = np.zeros(B.V, bool)
            np.logical_or.at(has_nan, c, nan)
            res = np.ldexp(sums.astype(np.float64).astype(np.float32), -self.shift).astype(np.float32)
            res[has_nan] = np.float32(np.nan)
            B.cost[s_c + p] = res

    # ---------------------------------------------------------------- proposals
    def init_bp(self, l):
        """init_bp (hbp_isp.cu:930-950): prior planes, costs of slots 0..2P-1, belief = cost,
        messages zero, the seeds (init_rand_seed_k, 570-591)."""
        B, Hh = self.lv[l], self.lv[l + 1]
        self.msg = np.zeros((B.E, K), np.float32)
        for p in range(P):                           # copy_prior_plane (56-73)
            pr = Hh.plane[p][B.clus]
            B.plane[p] = pr
            B.plane[p + P] = pr
            B.plane[p + 2 * P] = pr
        self.plane_cost(l, 0)
        self.plane_cost(l, P)
        B.belief = B.cost.copy()
        V, Vh = B.V, Hh.V
        self.s_neighbor[:V] = B.nb_pos[:V]
        self.s_ransac[:V] = B.pix_pos
        child_num = np.bincount(B.clus, minlength=Vh)
        child_pos = np.concatenate([[0], np.cumsum(child_num)[:-1]])
        self.s_child[:Vh] = child_pos
        self.child_num, self.child_pos = child_num, child_pos

    def ransac(self, l):
        """ransac_candidate_plane (hbp_isp.cu:592-640)."""
        B, W = self.lv[l], self.W
        for v in range(B.V):
            n, pos = int(B.pix_num[v]), int(B.pix_pos[v])
            s = int(self.s_ransac[v])
            for p in range(2 * P, 3 * P):
                s = lcg(s)
                p0 = umod(s, n)
                s = lcg(s)
                p1 = umod(s, n - 1)
                s = lcg(s)
                p2 = umod(s, n - 2)
                p0, p1, p2 = three_diff(p0, p1, p2, n)
                ids = [int(B.px[pos + q]) for q in (p0, p1, p2)]
                pts = [np.array([F(i % W), F(i // W), self.dwta[i], F(0)], np.float32) for i in ids]
                B.plane[p][v] = fit_plane(*pts)
                if p == 3 * P - 1:
                    B.point[v] = ids
            self.s_ransac[v] = s

    def spatial(self, l):
        """select_child_plane (705-726), neighbor_candidate_plane on the parent level (676-703)
        and set_candidate_plane_k (727-741)."""
        B, Hh = self.lv[l], self.lv[l + 1]
        for h in range(Hh.V):
            n, pos = int(self.child_num[h]), int(self.child_pos[h])
            s = int(self.s_child[h])
            for p in range(P):
                s = lcg(s)
                c = int(B.clus_sort[pos + umod(s, n)])
                Hh.plane[p + 2 * P][h] = B.plane[0][c]
            self.s_child[h] = s
        buf = np.zeros((P, Hh.V, 4), np.float32)
        for h in range(Hh.V):
            nb_num, nb_pos = int(Hh.nb_num[h]), int(Hh.nb_pos[h])
            s = int(self.s_neighbor[h])
            for p in range(P):
                s = lcg(s)
                if nb_num == 0:
                    t = h     # no edges (a one-node level): the node's own candidate
                else:
                    t = int(Hh.bd[nb_pos + umod(s, nb_num), 1])
                buf[p][h] = Hh.plane[p + 2 * P][t]
            self.s_neighbor[h] = s
        Hh.plane[2 * P:3 * P] = buf
        for p in range(2 * P, 3 * P):
            B.plane[p] = Hh.plane[p][B.clus]

    def random(self, l, max_dz, max_dn):
        """random_candidate_plane (hbp_isp.cu:641-674) with normalize_plane (495-503)."""
        B = self.lv[l]
        with np.errstate(all="ignore"):
            for v in range(B.V):
                s = int(self.s_ransac[v])
                x, y = B.mean[3][v], B.mean[4][v]
                for p in range(2 * P, 3 * P):
                    a, b, c = (F(q) for q in B.plane[p][v][:3])
                    z = F(F(F(a * x) + F(b * y)) + c)
                    s, dz = lcg_uniform(s, max_dz)
                    z = F(z + dz)
                    s, dnx = lcg_uniform(s, max_dn)
                    s, dny = lcg_uniform(s, max_dn)
                    s, dnz = lcg_uniform(s, max_dn)
                    norm = F(np.sqrt(F(F(F(a * a) + F(b * b)) + F(c * c))))
                    a, b, c = F(a / norm), F(b / norm), F(c / norm)
                    a, b, c = F(a + dnx), F(b + dny), F(c + dnz)
                    na = F(-a / c)
                    nb = F(-b / c)
                    nc = F(F(z - F(na * x)) - F(nb * y))
                    B.plane[p][v] = np.array([na, nb, nc, 0], np.float32)
                self.s_ransac[v] = s

    # ---------------------------------------------------------------- messages
    def belief_propagation(self, l):
        """passing_msg (hbp_isp.cu:1021-1047: passing_msg_k 292-304, smooth_msg_k 306-355,
        normalize_msg_k 273-290), aggregate_belief (1048-1058, 405-417), ping_pang_msg
        (1059-1066, 419-429), select_candidate_plane (1067-1079, 432-493)."""
        B, W = self.lv[l], self.W
        E = B.E
        msg = self.msg
        s_, t_ = B.bd[:, 0], B.bd[:, 1]
        with np.errstate(all="ignore"):
            for y in range(2 * P):                                    # passing_msg_k
                msg[:, y] = B.belief[y][t_] - msg[:, y]
            dist_pos = np.fmax(B.dc1, F(1.0))
            dist_color = np.fmax(B.dc0, F(0.01))
            c0 = (dist_color * B.bdl.astype(np.float32)).astype(np.float32)
            pts = np.concatenate([B.point[s_], B.point[t_]], axis=1)  # [E][6]
            fx = (pts % W).astype(np.float32)
            fy = (pts // W).astype(np.float32)
            msg_t = np.zeros((E, K), np.float32)
            for y in range(K):                                        # smooth_msg_k
                ps = B.plane[y][s_]
                acc = np.zeros(E, np.float32)
                minimum = np.full(E, FLT_MAX, np.float32)
                for p in range(2 * P):
                    if p == P:
                        acc = acc + minimum
                        minimum = np.full(E, FLT_MAX, np.float32)
                    pt = B.plane[p][t_]
                    dpx = ps[:, 0] - pt[:, 0]
                    dpy = ps[:, 1] - pt[:, 1]
                    dpz = ps[:, 2] - pt[:, 2]
                    phi = np.zeros(E, np.float32)
                    for k in range(6):
                        term = np.abs((dpx * fx[:, k] + dpy * fy[:, k]) + dpz)
                        phi = term if k == 0 else phi + term
                    # CUDA's min(float, float) is fminf: a NaN phi gives tau_s
                    phi = np.fmin((phi / F(3.0)) / dist_pos, self.tau_s)
                    m = msg[:, p] + c0 * phi
                    take = m < minimum
                    minimum = np.where(take, m, minimum)
                msg_t[:, y] = acc + minimum
            mean = np.zeros(E, np.float32)                            # normalize_msg_k
            for k in range(K):
                mean = mean + msg_t[:, k]
            mean = mean / F(K)
            for k in range(K):
                msg_t[:, k] = msg_t[:, k] - mean
            belief = B.cost.copy()                                    # aggregate_belief
            for e in range(E):
                s = s_[e]
                belief[:, s] = belief[:, s] + msg_t[e, :]
            B.belief = belief.astype(np.float32)
            for e in range(E):                                        # ping_pang_msg
                msg[B.rev[e], :] = msg_t[e, :]
            order = np.zeros((K, B.V), np.int64)                      # get_belief_order
            bel = B.belief
            for i in range(K - 1):
                for j in range(i + 1, K):
                    lt = bel[j] < bel[i]
                    order[i] += lt
                    order[j] += ~lt
            for v in range(B.V):                                      # select_belief
                nb = 0
                for k in range(K):
                    if order[k, v] < P:
                        B.plane[nb][v] = B.plane[k][v]
                        B.cost[nb][v] = B.cost[k][v]
                        B.belief[nb][v] = B.belief[k][v]
                        nb += 1
            for e in range(E):                                        # select_msg
                t = t_[e]
                nb = 0
                for k in range(K):
                    if order[k, t] < P:
                        msg[e, nb] = msg[e, k]
                        nb += 1

    # ---------------------------------------------------------------- driver
    def step(self, n=1 << 30):
        done = 0
        while done < n and self.next < len(self.sched):
            l, kind, dz, dn, first = self.sched[self.next]
            if first:
                self.init_bp(l)
            if kind == "ransac":
                self.ransac(l)
            elif kind == "spatial":
                self.spatial(l)
            else:
                self.random(l, dz, dn)
            self.plane_cost(l, 2 * P)
            self.belief_propagation(l)
            self.cur = l
            self.next += 1
            done += 1
        return done

    def state(self):
        B = self.lv[self.cur]
        return dict(level=self.cur, V=B.V, E=B.E, planes=B.plane.copy(), cost=B.cost.copy(),
                    belief=B.belief.copy(), msg=self.msg[:, :2 * P].copy())

    def map(self):
        """map_msg (hbp_isp.cu:1081-1105: map_msg_k 783-794, min_belief_to_plane 796-817,
        plane_to_label 819-841)."""
        B = self.lv[self.min_level]
        W = self.W
        best = np.zeros((B.V, 4), np.float32)
        for v in range(B.V):
            mb = FLT_MAX
            for p in range(P):
                mb = amin_fold(mb, B.belief[p][v])
            mp = 0
            for p in range(P):
                if B.belief[p][v] == mb:
                    mp = p
                    break
            best[v] = B.plane[mp][v]
        c = B.px_clus
        x = (B.px % W).astype(np.float32)
        y = (B.px // W).astype(np.float32)
        pl = best[c]
        with np.errstate(all="ignore"):
            d = (pl[:, 0] * x + pl[:, 1] * y) + pl[:, 2]
            d = np.fmin(np.fmax(d, F(0)), F(self.D - 1))
        disp = np.zeros(self.N, np.float32)
        plane3 = np.zeros((self.N, 3), np.float32)
        disp[B.px] = d
        plane3[B.px] = pl[:, :3]
        return disp.reshape(self.H, W), plane3.reshape(self.H, W, 3)

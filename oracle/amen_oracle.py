"""CPU oracle for the AMEn linear-solver path -- TEST INFRASTRUCTURE ONLY.

A numpy FP64 restatement of the reference's native AMEn solver
(cpp/amen_solve.h:122-549, cpp/gmres.h:15-161, cpp/ortho.h:16-73) and of the sweeps around
it (preconditioners, amen_divide, the AMEn TT products, the DMRG matvec, TT rounding),
written to be read line against line with them. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker / baseline; the
product path (torchtt_amd) never does.

Pinning (see tests/test_oracle.py):
  * amen_solve: against the reference's own C++ build (oracle/_ref) on identical
    seeded inputs, against the closed-form Laplacian solution (laplace_exact.py),
    and by the TT residual ||Ax-b||/||b|| of its result;
  * gmres: against numpy.linalg.solve on small dense systems, and the Arnoldi
    relation A Q_k = Q_{k+1} Hbar_k;
  * local_product / phi_*: against brute-force dense contractions and the
    invariants phi_d = x^T A x and phib_d = <b, x>;
  * rank_chop / round_tt: against the closed-form tail rule and dense SVD;
  * prec_blocks / prec_apply (cpp/matvecs.h:153-194): exact inverse of a block-diagonal operator;
  * amen_divide (torchtt/_division.py:76-398): the dense quotient and the rank-2 truncation floor;
  * amen_mv / kron_eye_operator (torchtt/_amen.py:55-413): exact products, zero and identity
    operands, amen_mm through A (x) I_N;
  * dmrg_mv (cpp/dmrg_mv.h:23-230, tests/test_dmrg_oracle.py): dense products, the reference
    test's structure, identity and zero operands, the final-rank rule; randn_stream against the
    published SplitMix64 outputs and the N(0,1) moments.
The reference's Python twin (torchtt/solvers.py:290-688) differs in three places
(GMRES iterations m+1, truncation start r = full-1, zero-norm guard 1.0); this
module follows the C++ version.
"""
import numpy as np

__all__ = ["local_product", "phi_fwd_A", "phi_bck_A", "phi_fwd_rhs", "phi_bck_rhs",
           "gmres", "amen_solve", "rl_orthogonal", "rank_chop", "round_tt", "tt_full",
           "tt_matvec", "tt_add", "tt_norm", "tt_residual", "prec_blocks", "prec_apply", "diag_operator",
           "amen_divide", "local_AB", "amen_mv", "kron_eye_operator", "randn_stream", "dmrg_mv"]


# --------------------------------------------------------------------------- contractions
def local_product(phi_right, phi_left, core_a, core):
    """w[l,m,L] = sum Phi_L[l,s,r] A[s,m,n,S] Phi_R[L,S,R] x[r,n,R]  (amen_solve.h:17-27)."""
    t1 = np.tensordot(core, phi_left, axes=([0], [2]))           # n R l s
    t2 = np.tensordot(t1, core_a, axes=([0, 3], [2, 0]))         # R l m S
    return np.tensordot(t2, phi_right, axes=([0, 3], [2, 1]))    # l m L


def prec_blocks(mode, phi_left, core_a, phi_right):
    """Local preconditioner blocks (AMENsolveMV::setter, matvecs.h:161-174).

    mode 'c' (C_PREC): J[l, L] = inv_{m,n}( sum_{s,S} PhiL[l,s,l] A[s,m,n,S] PhiR[L,S,L] ), shape r1 x r2 x n x n.
    mode 'r' (R_PREC): J[l] = inv_{(m,L),(n,R)}( sum_{s,S} PhiL[l,s,l] A[s,m,n,S] PhiR[L,S,R] ),
    shape r1 x n x r2 x n x r2.
    """
    jl = np.tensordot(np.diagonal(phi_left, 0, 0, 2), core_a, axes=([0], [0]))   # d m n S
    if mode == 'c':
        jr = np.diagonal(phi_right, 0, 0, 2)                                       # S D
        j = np.tensordot(jl, jr, axes=([3], [0])).transpose(0, 3, 1, 2)            # d D m n
        return np.linalg.inv(j)
    jt = np.tensordot(jl, phi_right, axes=([3], [1])).transpose(0, 1, 3, 2, 4)      # d m L n R
    sh = jt.shape
    j = jt.reshape(sh[0], sh[1] * sh[2], sh[3] * sh[4])
    return np.linalg.inv(j).reshape(sh)


def prec_apply(mode, j, x):
    """y = P x (AMENsolveMV::apply_prec, matvecs.h:177-193); x is r1 x n x r2."""
    if mode == 'c':
        return np.einsum('rRmn,rnR->rmR', j, x)
    return np.einsum('rmLnR,rnR->rmL', j, x)


def phi_bck_A(phi_now, core_left, core_a, core_right):
    """Phi_prev[l,s,r] = sum xL[l,M,L] Phi[L,S,R] A[s,M,N,S] xR[r,N,R]  (amen_solve.h:38-47)."""
    p = np.tensordot(core_left, phi_now, axes=([2], [0]))        # l M S R
    p = np.tensordot(p, core_a, axes=([1, 2], [1, 3]))           # l R s N
    return np.tensordot(p, core_right, axes=([1, 3], [2, 1]))    # l s r


def phi_fwd_A(phi_now, core_left, core_a, core_right):
    """Phi_next[L,S,R] = sum Phi[l,s,r] xL[l,M,L] A[s,M,N,S] xR[r,N,R]  (amen_solve.h:58-68)."""
    p = np.tensordot(core_left, phi_now, axes=([0], [0]))        # M L s r
    p = np.tensordot(p, core_a, axes=([0, 2], [1, 0]))           # L r N S
    return np.tensordot(p, core_right, axes=([1, 2], [0, 1]))    # L S R


def phi_bck_rhs(phi_now, core_b, core):
    """Phib_prev[b,r] = sum b[b,n,B] Phib[B,R] x[r,n,R]  (amen_solve.h:78-83)."""
    p = np.tensordot(core_b, phi_now, axes=([2], [0]))           # b n R
    return np.tensordot(p, core, axes=([1, 2], [1, 2]))          # b r


def phi_fwd_rhs(phi_now, core_b, core):
    """Phib_next[B,R] = sum Phib[b,r] b[b,n,B] x[r,n,R]  (amen_solve.h:93-98)."""
    p = np.tensordot(phi_now, core_b, axes=([0], [0]))           # r n B
    return np.tensordot(p, core, axes=([0, 1], [0, 1]))          # B R


# --------------------------------------------------------------------------- GMRES
def _gmres_single(matvec, rhs, x0, iters, threshold):
    """One GMRES(iters) cycle with MGS and Givens rotations (gmres.h:34-144)."""
    r = rhs - matvec(x0)
    b_norm = np.linalg.norm(rhs)
    r_norm = np.linalg.norm(r)
    if r_norm <= 0:
        return x0.copy(), 1, 0
    Q = [r / r_norm]
    H = np.zeros((iters + 1, iters))
    beta = np.zeros(iters + 1)
    beta[0] = r_norm
    cs = np.zeros(iters)
    sn = np.zeros(iters)
    flag = 0
    k = 0
    for k in range(iters):
        q = matvec(Q[k])
        for i in range(k + 1):                      # modified Gram-Schmidt (:98-101)
            H[i, k] = q @ Q[i]
            q = q - H[i, k] * Q[i]
        h = np.linalg.norm(q)
        with np.errstate(divide="ignore", invalid="ignore"):
            q = q / h
        H[k + 1, k] = h
        Q.append(q)
        hcol = H[:k + 2, k].copy()                  # Givens (:20-31, :113-116)
        for i in range(k):
            t = cs[i] * hcol[i] + sn[i] * hcol[i + 1]
            hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1]
            hcol[i] = t
        den = np.sqrt(hcol[k] ** 2 + hcol[k + 1] ** 2)
        c, s = hcol[k] / den, hcol[k + 1] / den
        hcol[k] = c * hcol[k] + s * hcol[k + 1]
        hcol[k + 1] = 0.0
        H[:k + 2, k] = hcol
        cs[k], sn[k] = c, s
        beta[k + 1] = -sn[k] * beta[k]              # (:120-122)
        beta[k] = cs[k] * beta[k]
        if abs(beta[k + 1]) / b_norm <= threshold:
            flag = 1
            break
    k = k if k < iters else iters - 1               # (:131)
    y = np.linalg.solve(H[:k + 1, :k + 1], beta[:k + 1])
    sol = x0.copy()
    for i in range(k + 1):
        sol = sol + Q[i] * y[i]
    return sol, flag, k + 1


def gmres(matvec, rhs, x0, max_iters, threshold, resets):
    """Restarted GMRES (gmres.h:146-161). Returns (solution, flag, iterations)."""
    nit, flag = 0, 0
    xs = x0
    sol = x0
    for _ in range(resets):
        sol, flag, nowit = _gmres_single(matvec, rhs, xs, max_iters, threshold)
        nit += nowit
        if flag == 1:
            break
        xs = sol.copy()
    return sol, flag, nit


# --------------------------------------------------------------------------- QR / rank tools
def _qr(m):
    q, r = np.linalg.qr(m, mode="reduced")
    return q, r


def rl_orthogonal(cores):
    """Right-to-left orthogonalisation in place (ortho.h:46-73). Returns the new ranks."""
    d = len(cores)
    for i in range(d - 1, 0, -1):
        r0, n, r1 = cores[i].shape
        q, r = _qr(cores[i].reshape(r0, n * r1).T)
        rn = r.shape[0]
        cores[i] = q.T.reshape(rn, n, r1)
        rp, np_, _ = cores[i - 1].shape
        cores[i - 1] = (cores[i - 1].reshape(-1, r0) @ r.T).reshape(rp, np_, rn)
    return [c.shape[0] for c in cores] + [1]


def rank_chop(s, eps):
    """Smallest rank whose discarded tail has 2-norm below eps (ortho.h:16-44)."""
    n = s.shape[0]
    r = n - 1
    if np.linalg.norm(s) == 0.0:
        return 1
    if eps <= 0.0:
        return r
    while r > 0:
        if np.sum(s[r:] ** 2) >= eps * eps:
            break
        r -= 1
    r += 1
    return max(r, 1)


def round_tt(cores, epsilon, rmax):
    """TT rounding (compression.h:3-41): left-to-right QR then right-to-left SVD."""
    cores = [c.copy() for c in cores]
    d = len(cores)
    if d <= 1:
        return cores
    for i in range(d - 1):                                  # lr_orthogonal (ortho.h:77-101)
        r0, n, r1 = cores[i].shape
        q, r = _qr(cores[i].reshape(r0 * n, r1))
        cores[i] = q.reshape(r0, n, q.shape[1])
        rn, nn, rr = cores[i + 1].shape
        cores[i + 1] = (r @ cores[i + 1].reshape(rn, -1)).reshape(q.shape[1], nn, rr)
    eps = epsilon / np.sqrt(d - 1.0)
    core_now = cores[d - 1].reshape(cores[d - 1].shape[0], -1)
    for i in range(d - 1, 0, -1):
        r0, n, r1 = cores[i - 1].shape
        core_next = cores[i - 1].reshape(r0 * n, r1)
        u, s, vt = np.linalg.svd(core_now, full_matrices=False)
        rc = rank_chop(s, eps * np.linalg.norm(s))
        rnew = min(rmax, rc)
        us = u[:, :rnew] @ np.diag(s[:rnew])
        core_next = core_next @ us
        cores[i] = vt[:rnew, :].reshape(rnew, cores[i].shape[1], cores[i].shape[2])
        cores[i - 1] = core_next.reshape(r0, n, rnew)
        core_now = core_next.reshape(r0, -1)
    return cores


# --------------------------------------------------------------------------- AMEn
def amen_solve(A, b, x0=None, z0=None, nswp=22, eps=1e-10, rmax=32768, max_full=500,
               kickrank=4, kick2=0, local_iterations=40, resets=2, rng=None, trace=None,
               preconditioner=None, python_rules=False, rank_caps=None):
    """AMEn solve following cpp/amen_solve.h:122-549 (FP64; preconditioner None, 'c' or 'r'
    right-preconditions the iterative local solve as in amen_solve.h:346-356).

    A: operator cores rA x n x n x rA'; b: rhs cores; x0: optional initial cores;
    z0: initial residual-proxy cores rz x n x rz' (normal random; the reference draws
    them with torch.randn, so a test passes the same draws). `trace`, if a dict, receives
    per-sweep ranks, max_res and GMRES iteration counts, and per core (sweep, core, res_old,
    res_new, rank). python_rules: the truncation loop and zero-norm guard of the reference's
    Python solvers (see amen_divide); rank_caps: d+1 per-interface rank caps instead of rmax.
    """
    rng = np.random.default_rng(0) if rng is None else rng
    d = len(A)
    N = [a.shape[2] for a in A]
    if x0 is None:
        x = [np.ones((1, n, 1)) for n in N]                     # :161-164
    else:
        x = [c.copy() for c in x0]
    rx = [1] + [c.shape[2] for c in x]
    rz = [1] + [kickrank + kick2] * (d - 1) + [1]               # :169-173
    if z0 is None:
        z = [rng.standard_normal((rz[i], N[i], rz[i + 1])) for i in range(d)]
    else:
        z = [c.copy() for c in z0]
    rz = rl_orthogonal(z)                                        # :178
    rz = [1] + rz[1:-1] + [1]
    one3, one2 = np.ones((1, 1, 1)), np.ones((1, 1))
    Phiz = [one3] + [None] * (d - 1) + [one3]
    Phiz_b = [one2] + [None] * (d - 1) + [one2]
    Phis = [one3] + [None] * (d - 1) + [one3]
    Phis_b = [one2] + [None] * (d - 1) + [one2]
    normA = np.ones(max(d - 1, 0))
    normb = np.ones(max(d - 1, 0))
    normx = np.ones(max(d - 1, 0))
    nrmsc = 1.0
    damp = 2.0
    # a zero interface norm divides by 0 in the C++ solver (amen_solve.h:275, 280) and by 1 in the
    # Python solvers (_division.py:166-173)
    zero_norm = 1.0 if python_rules else 0.0
    last = False
    if trace is not None:
        trace.setdefault("ranks", [])
        trace.setdefault("max_res", [])
        trace.setdefault("gmres_iters", [])
    swp = 0
    for swp in range(nswp):
        for k in range(d - 1, 0, -1):                            # :222-292
            if not last:
                if swp > 0:
                    czA = local_product(Phiz[k + 1], Phiz[k], A[k], x[k])
                    czy = np.tensordot(Phiz_b[k], b[k], axes=([0], [0]))
                    czy = np.tensordot(czy, Phiz_b[k + 1], axes=([2], [0]))
                    czy = czy * nrmsc - czA
                    _, _, vt = np.linalg.svd(czy.reshape(czy.shape[0], -1), full_matrices=False)
                    temp = min(kickrank, vt.shape[0])
                    cz_new = vt[:temp, :].T
                    if k < d - 1 and kick2 > 0:
                        cz_new = np.concatenate([cz_new, rng.standard_normal((cz_new.shape[0], kick2))], 1)
                else:
                    cz_new = z[k].reshape(rz[k], -1).T
                qz, _ = _qr(cz_new)
                rz[k] = qz.shape[1]
                z[k] = qz.T.reshape(rz[k], N[k], rz[k + 1])
            if swp > 0:
                nrmsc = nrmsc * normA[k - 1] * normx[k - 1] / normb[k - 1]
            core = x[k].reshape(rx[k], N[k] * rx[k + 1]).T       # :252-267
            q, r = _qr(core)
            core_prev = np.tensordot(x[k - 1], r.T, axes=([2], [0]))
            rx[k] = q.shape[1]
            current_norm = np.linalg.norm(core_prev)
            if current_norm > 0:
                core_prev = core_prev / current_norm
            else:
                current_norm = 1.0
            normx[k - 1] *= current_norm
            x[k] = q.T.reshape(rx[k], N[k], rx[k + 1]).copy()
            x[k - 1] = core_prev.copy()
            Phis[k] = phi_bck_A(Phis[k + 1], x[k], A[k], x[k])   # :270-285
            Phis_b[k] = phi_bck_rhs(Phis_b[k + 1], b[k], x[k])
            nrm = np.linalg.norm(Phis[k])
            nrm = nrm if nrm > 0 else zero_norm
            normA[k - 1] = nrm
            Phis[k] = Phis[k] / nrm
            nrm = np.linalg.norm(Phis_b[k])
            nrm = nrm if nrm > 0 else zero_norm
            normb[k - 1] = nrm
            Phis_b[k] = Phis_b[k] / nrm
            nrmsc = nrmsc * normb[k - 1] / (normA[k - 1] * normx[k - 1])
            if not last:                                          # :288-291
                Phiz[k] = phi_bck_A(Phiz[k + 1], z[k], A[k], x[k]) / normA[k - 1]
                Phiz_b[k] = phi_bck_rhs(Phiz_b[k + 1], b[k], z[k]) / normb[k - 1]
        max_res = 0.0
        gm_its = []
        for k in range(d):                                        # :296-512
            previous_solution = x[k].reshape(-1)
            rhs = np.tensordot(Phis_b[k], b[k] * nrmsc, axes=([0], [0]))
            rhs = np.tensordot(rhs, Phis_b[k + 1], axes=([2], [0])).reshape(-1)
            norm_rhs = np.linalg.norm(rhs)
            real_tol = (eps / np.sqrt(d)) / damp
            shape_now = (rx[k], N[k], rx[k + 1])
            use_full = rx[k] * N[k] * rx[k + 1] < max_full
            pl, pr, ak = Phis[k], Phis[k + 1], A[k]

            def op(v, pl=pl, pr=pr, ak=ak, shape_now=shape_now):
                return local_product(pr, pl, ak, v.reshape(shape_now)).reshape(-1)

            if use_full:                                          # :319-330
                Bp = np.tensordot(ak, pr, axes=([3], [1]))        # s m n L R
                Bp = np.tensordot(pl, Bp, axes=([1], [0]))        # l r m n L R
                nl = rx[k] * N[k] * rx[k + 1]
                B = Bp.transpose(0, 2, 4, 1, 3, 5).reshape(nl, nl)
                solution_now = np.linalg.solve(B, rhs)
                res_old = np.linalg.norm(B @ previous_solution - rhs) / norm_rhs
                res_new = np.linalg.norm(B @ solution_now - rhs) / norm_rhs

                def apply(v, B=B):
                    return B @ v
                gm_its.append(0)
            else:                                                 # :331-365
                eps_local = real_tol * norm_rhs
                drhs = rhs - op(previous_solution)
                eps_local /= np.linalg.norm(drhs)
                ps = 0.0 * previous_solution
                if preconditioner is None:
                    solution_now, flag, nit = gmres(op, drhs, ps, local_iterations, eps_local, resets)
                else:                                             # matvecs.h:198-212, amen_solve.h:353-356
                    J = prec_blocks(preconditioner, pl, ak, pr)

                    def pop(v, J=J, shape_now=shape_now):
                        return op(prec_apply(preconditioner, J, v.reshape(shape_now)).reshape(-1))
                    solution_now, flag, nit = gmres(pop, drhs, ps, local_iterations, eps_local, resets)
                    solution_now = prec_apply(preconditioner, J, solution_now.reshape(shape_now)).reshape(-1)
                gm_its.append(nit)
                solution_now = solution_now + previous_solution
                res_old = np.linalg.norm(op(previous_solution) - rhs) / norm_rhs
                res_new = np.linalg.norm(op(solution_now) - rhs) / norm_rhs
                apply = op
            max_res = max(max_res, res_old)
            solution_now = solution_now.reshape(rx[k] * N[k], rx[k + 1])
            if k < d - 1:                                         # :382-404
                u, s, vt = np.linalg.svd(solution_now, full_matrices=False)
                v = vt
                cap = rmax if rank_caps is None else rank_caps[k + 1]
                if python_rules:                                  # _division.py:283-296
                    r = 0
                    for r in range(u.shape[1] - 1, 0, -1):
                        sol = (u[:, :r] * s[:r]) @ v[:r, :]
                        res = np.linalg.norm(apply(sol.reshape(-1)) - rhs) / norm_rhs
                        if res > max(real_tol * damp, res_new):
                            break
                    r += 1
                    r = min(r, s.shape[0], cap)
                else:
                    r = u.shape[1]
                    while r > 0:
                        sol = (u[:, :r] * s[:r]) @ v[:r, :]
                        res = np.linalg.norm(apply(sol.reshape(-1)) - rhs) / norm_rhs
                        if res > max(res_new, real_tol * damp):
                            break
                        r -= 1
                    r += 1
                    r = r if (r < u.shape[1] and r < cap) else min(u.shape[1], cap)
            else:                                                 # :406-410
                u, v = _qr(solution_now)
                r = u.shape[1]
                s = np.ones(r)
            if trace is not None:
                trace.setdefault("cores", []).append((swp, k, res_old, res_new, int(r)))
            u = u[:, :r]                                          # :413-416
            v = (np.diag(s[:r]) @ v[:r, :]).T
            if not last:                                          # :418-442
                tmp = (u @ v.T).reshape(rx[k], N[k], rx[k + 1])
                czA = local_product(Phiz[k + 1], Phiz[k], A[k], tmp)
                czy = np.tensordot(Phiz_b[k], nrmsc * b[k], axes=([0], [0]))
                czy = np.tensordot(czy, Phiz_b[k + 1], axes=([2], [0]))
                tmp = (czy - czA).reshape(rz[k] * N[k], rz[k + 1])
                uz, _, _ = np.linalg.svd(tmp, full_matrices=False)
                rtmp = min(kickrank, uz.shape[1])
                tmp = uz[:, :rtmp]
                if k < d - 1 and kick2 > 0:
                    tmp = np.concatenate([tmp, rng.standard_normal((tmp.shape[0], kick2))], 1)
                qz, _ = _qr(tmp)
                rz[k + 1] = qz.shape[1]
                z[k] = qz.reshape(rz[k], N[k], rz[k + 1]).copy()
            if k < d - 1:                                         # :443-504
                if not last:
                    tmp = (u @ v.T).reshape(rx[k], N[k], rx[k + 1])
                    left_res = local_product(Phiz[k + 1], Phis[k], A[k], tmp)
                    left_b = np.tensordot(Phis_b[k], b[k] * nrmsc, axes=([0], [0]))
                    left_b = np.tensordot(left_b, Phiz_b[k + 1], axes=([2], [0]))
                    uk = (left_b - left_res).reshape(u.shape[0], -1)
                    uk = np.concatenate([u, uk], 1)
                    r_add = left_res.shape[2]
                    u, Rmat = _qr(uk)
                    v = np.concatenate([v, np.zeros((rx[k + 1], r_add))], 1)
                    v = v @ Rmat.T
                r = u.shape[1]
                v = np.tensordot(v, x[k + 1], axes=([0], [0]))
                nrmsc = nrmsc * normA[k] * normx[k] / normb[k]
                norm_now = np.linalg.norm(v)
                if norm_now > 0:
                    v = v / norm_now
                else:
                    norm_now = 1.0
                normx[k] *= norm_now
                x[k] = u.reshape(rx[k], N[k], r).copy()
                x[k + 1] = v.reshape(r, N[k + 1], rx[k + 2]).copy()
                rx[k + 1] = r
                Phis[k + 1] = phi_fwd_A(Phis[k], x[k], A[k], x[k])
                Phis_b[k + 1] = phi_fwd_rhs(Phis_b[k], b[k], x[k])
                nrm = np.linalg.norm(Phis[k + 1])
                nrm = nrm if nrm > 0 else 1.0
                normA[k] = nrm
                Phis[k + 1] = Phis[k + 1] / nrm
                nrm = np.linalg.norm(Phis_b[k + 1])
                nrm = nrm if nrm > 0 else 1.0
                normb[k] = nrm
                Phis_b[k + 1] = Phis_b[k + 1] / nrm
                nrmsc = nrmsc * normb[k] / (normA[k] * normx[k])
                if not last:
                    Phiz[k + 1] = phi_fwd_A(Phiz[k], z[k], A[k], x[k]) / normA[k]
                    Phiz_b[k + 1] = phi_fwd_rhs(Phiz_b[k], b[k], z[k]) / normb[k]
            else:
                usv = (u * s[:r]) @ v[:r, :].T
                x[k] = usv.reshape(rx[k], N[k], rx[k + 1])
        if trace is not None:
            trace["ranks"].append(list(rx))
            trace["max_res"].append(max_res)
            trace["gmres_iters"].append(gm_its)
        if last:
            break
        if max_res < eps:
            last = True
    if trace is not None:
        trace["sweeps"] = swp + 1
    norm_x = np.exp(np.sum(np.log(normx)) / d) if d > 1 else 1.0   # :534-541
    return [c * norm_x for c in x]


def diag_operator(a):
    """TT-matrix cores of diag(a) for a TT tensor a: D[s,m,n,S] = a[s,m,S] if m == n else 0 (the
    operator of the reference's LinearOp, _division.py:30-75, 'lsr,smS,LSR,rmR->lmL')."""
    out = []
    for c in a:
        r0, n, r1 = c.shape
        D = np.zeros((r0, n, n, r1))
        idx = np.arange(n)
        D[:, idx, idx, :] = c
        out.append(D)
    return out


def amen_divide(a, b, x0=None, z0=None, nswp=22, eps=1e-10, rmax=100, max_full=500, kickrank=4,
                kick2=0, local_iterations=40, resets=2, rng=None, trace=None, preconditioner=None):
    """x = b ./ a by AMEn, following torchtt/_division.py:76-398 (the reference's Python solver):
    the sweep of amen_solve on the operator diag(a), started from all-ones cores of ranks
    1, 2, ..., 2, 1 (:86-88) with per-interface rank caps (:94-95), GMRES restarts of
    local_iterations + 1 steps (:253), the Python truncation loop (:283-296) and zero-norm
    guard 1 (:166-173)."""
    d = len(a)
    N = [c.shape[1] for c in a]
    caps = [1] + (d - 1) * [int(rmax)] + [1] if np.isscalar(rmax) else [int(v) for v in rmax]
    if x0 is None:
        rx = [1] + (d - 1) * [2] + [1]
        x0 = [np.ones((rx[k], N[k], rx[k + 1])) for k in range(d)]
    return amen_solve(diag_operator(a), b, x0=x0, z0=z0, nswp=nswp, eps=eps, rmax=max(caps), max_full=max_full,
                      kickrank=kickrank, kick2=kick2, local_iterations=local_iterations + 1, resets=resets,
                      rng=rng, trace=trace, preconditioner=preconditioner, python_rules=True, rank_caps=caps)


# --------------------------------------------------------------------------- AMEn products
def local_AB(PhiL, PhiR, a, b):
    """w[r,m,R] = sum PhiL[r,a,b] A[a,m,k,A] b[b,k,B] PhiR[R,A,B] (_amen.py:36-52, vector b)."""
    return np.einsum("rab,amkA,bkB,RAB->rmR", PhiL, a, b, PhiR, optimize=True)


def amen_mv(A, b, x0=None, z0=None, nswp=22, eps=1e-10, rmax=1024, kickrank=4, kick2=0, rng=None,
            trace=None):
    """y ~= A b by AMEn (torchtt/_amen.py:55-101 -> _amen_mm_python :126-413, vector case): the
    solution cores are recomputed from the projected product, truncated by rank_chop at
    ||w|| eps / d^(1.5, or 0.5 in the last sweep) (:297-300), enriched with the residual proxy z;
    the sweep stops one sweep after max dx < eps (:391-397; a NaN dx, 0/0 for a zero product,
    never stops it). x0 defaults to zero cores of rank 1 (:134-137)."""
    rng = np.random.default_rng(0) if rng is None else rng
    d = len(A)
    P = [a.shape[1] for a in A]
    if x0 is None:
        x = [np.zeros((1, p, 1)) for p in P]
    else:
        x = [c.copy() for c in x0]
    rx = [1] + [c.shape[2] for c in x]
    caps = [1] + (d - 1) * [int(rmax)] + [1] if np.isscalar(rmax) else [int(v) for v in rmax]
    rz = [1] + (d - 1) * [kickrank + kick2] + [1]
    z = [c.copy() for c in z0] if z0 is not None else [rng.standard_normal((rz[k], P[k], rz[k + 1])) for k in range(d)]
    rz = rl_orthogonal(z)                                          # :146-153
    rz = [1] + rz[1:-1] + [1]
    one2, one3 = np.ones((1, 1)), np.ones((1, 1, 1))
    Phiz = [one2] + [None] * (d - 1) + [one2]
    Phiz_rhs = [one3] + [None] * (d - 1) + [one3]
    Phis = [one2] + [None] * (d - 1) + [one2]
    Phis_rhs = [one3] + [None] * (d - 1) + [one3]
    normA = np.ones(max(d - 1, 0))                                 # stays 1 (:201-205 disabled)
    normb = np.ones(max(d - 1, 0))
    normx = np.ones(max(d - 1, 0))
    nrmsc = 1.0
    last = False
    if trace is not None:
        trace.setdefault("ranks", [])
        trace.setdefault("max_dx", [])
    bx = lambda Phi, l, r: np.einsum("LR,lpL,rpR->lr", Phi, l, r)          # _compute_phi_bck_x
    fx = lambda Phi, l, r: np.einsum("lr,lpL,rpR->LR", Phi, l, r)          # _compute_phi_fwd_x
    bAB = lambda Phi, a, bb, c: np.einsum("RAB,amkA,bkB,rmR->rab", Phi, a, bb, c, optimize=True)
    fAB = lambda Phi, a, bb, c: np.einsum("rab,amkA,bkB,rmR->RAB", Phi, a, bb, c, optimize=True)
    proj = lambda L, c, R: np.einsum("zr,rpR,ZR->zpZ", L, c, R)
    swp = 0
    for swp in range(nswp):
        for k in range(d - 1, 0, -1):                              # :199-262
            if not last:
                if swp > 0:
                    cz_new = local_AB(Phiz_rhs[k], Phiz_rhs[k + 1], A[k], b[k]) * nrmsc - proj(Phiz[k], x[k], Phiz[k + 1])
                    _, _, vt = np.linalg.svd(cz_new.reshape(cz_new.shape[0], -1), full_matrices=False)
                    cz_new = vt[:min(kickrank, vt.shape[0]), :].T
                    if k < d - 1 and kick2 > 0:
                        cz_new = np.concatenate([cz_new, rng.standard_normal((cz_new.shape[0], kick2))], 1)
                else:
                    cz_new = z[k].reshape(rz[k], -1).T
                qz, _ = _qr(cz_new)
                rz[k] = qz.shape[1]
                z[k] = qz.T.reshape(rz[k], P[k], rz[k + 1])
            if swp > 0:
                nrmsc = nrmsc * normA[k - 1] * normx[k - 1] / normb[k - 1]
            q, r = _qr(x[k].reshape(rx[k], -1).T)
            core_prev = np.tensordot(x[k - 1], r.T, axes=([2], [0]))
            rx[k] = q.shape[1]
            cn = np.linalg.norm(core_prev)
            if cn > 0:
                core_prev = core_prev / cn
            else:
                cn = 1.0
            normx[k - 1] *= cn
            x[k] = q.T.reshape(rx[k], P[k], rx[k + 1])
            x[k - 1] = core_prev
            Phis[k] = bx(Phis[k + 1], x[k], x[k])
            Phis_rhs[k] = bAB(Phis_rhs[k + 1], A[k], b[k], x[k])
            nrm = np.linalg.norm(Phis_rhs[k])
            nrm = nrm if nrm > 0 else 1.0
            normb[k - 1] = nrm
            Phis_rhs[k] = Phis_rhs[k] / nrm
            nrmsc = nrmsc * normb[k - 1] / (normA[k - 1] * normx[k - 1])
            if not last:
                Phiz[k] = bx(Phiz[k + 1], z[k], x[k]) / normA[k - 1]
                Phiz_rhs[k] = bAB(Phiz_rhs[k + 1], A[k], b[k], z[k]) / normb[k - 1]
        max_dx = 0.0
        for k in range(d):                                         # :265-388
            prev = x[k]
            sol = local_AB(Phis_rhs[k], Phis_rhs[k + 1], A[k], b[k]) * nrmsc
            nsol = np.linalg.norm(sol)
            with np.errstate(invalid="ignore", divide="ignore"):
                dx = np.linalg.norm(sol - prev) / nsol
            max_dx = max(dx, max_dx)                               # Python max: a NaN dx stays
            sol = sol.reshape(rx[k] * P[k], rx[k + 1])
            if k < d - 1:
                u, s, vt = np.linalg.svd(sol, full_matrices=False)
                r = rank_chop(s, nsol * eps / (d ** (0.5 if last else 1.5)))
                r = min(r, s.shape[0], caps[k + 1])
            else:
                u, vt = _qr(sol)
                r = u.shape[1]
                s = np.ones(r)
            u = u[:, :r]
            v = (np.diag(s[:r]) @ vt[:r, :]).T
            if not last:
                cz_new = (local_AB(Phiz_rhs[k], Phiz_rhs[k + 1], A[k], b[k]) * nrmsc
                          - proj(Phiz[k], (u @ v.T).reshape(rx[k], P[k], rx[k + 1]), Phiz[k + 1]))
                uz, _, _ = np.linalg.svd(cz_new.reshape(rz[k] * P[k], rz[k + 1]), full_matrices=False)
                cz_new = uz[:, :min(kickrank, uz.shape[1])]
                if k < d - 1 and kick2 > 0:
                    cz_new = np.concatenate([cz_new, rng.standard_normal((cz_new.shape[0], kick2))], 1)
                qz, _ = _qr(cz_new)
                rz[k + 1] = qz.shape[1]
                z[k] = qz.reshape(rz[k], P[k], rz[k + 1])
            if k < d - 1:
                if not last:
                    uk = (local_AB(Phis_rhs[k], Phiz_rhs[k + 1], A[k], b[k]) * nrmsc
                          - proj(Phis[k], (u @ v.T).reshape(rx[k], P[k], rx[k + 1]), Phiz[k + 1]))
                    u, Rm = _qr(np.concatenate([u, uk.reshape(u.shape[0], -1)], 1))
                    v = np.concatenate([v, np.zeros((rx[k + 1], uk.shape[-1]))], 1) @ Rm.T
                r = u.shape[1]
                v = np.tensordot(v, x[k + 1], axes=([0], [0]))
                nrmsc = nrmsc * normA[k] * normx[k] / normb[k]
                nn = np.linalg.norm(v)
                if nn > 0:
                    v = v / nn
                else:
                    nn = 1.0
                normx[k] *= nn
                x[k] = u.reshape(rx[k], P[k], r)
                x[k + 1] = v.reshape(r, P[k + 1], rx[k + 2])
                rx[k + 1] = r
                Phis[k + 1] = fx(Phis[k], x[k], x[k])
                Phis_rhs[k + 1] = fAB(Phis_rhs[k], A[k], b[k], x[k])
                nrm = np.linalg.norm(Phis_rhs[k + 1])
                nrm = nrm if nrm > 0 else 1.0
                normb[k] = nrm
                Phis_rhs[k + 1] = Phis_rhs[k + 1] / nrm
                nrmsc = nrmsc * normb[k] / (normA[k] * normx[k])
                if not last:
                    Phiz[k + 1] = fx(Phiz[k], z[k], x[k]) / normA[k]
                    Phiz_rhs[k + 1] = fAB(Phiz_rhs[k], A[k], b[k], z[k]) / normb[k]
            else:
                x[k] = (u @ np.diag(s[:r]) @ v[:r, :].T).reshape(rx[k], P[k], rx[k + 1])
        if trace is not None:
            trace["ranks"].append(list(rx))
            trace["max_dx"].append(max_dx)
        if last:
            break
        if max_dx < eps:
            last = True
    if trace is not None:
        trace["sweeps"] = swp + 1
    nx = np.exp(np.sum(np.log(normx)) / d) if d > 1 else 1.0
    return [c * nx for c in x]


def kron_eye_operator(A, N):
    """Cores of A (x) I_N per mode, so that amen_mm(A, B) = amen_mv(A (x) I, vec B) with x cores
    indexed [r, (m, n), r'] as _amen_mm_python keeps them (:134-137, :393-411)."""
    out = []
    for a, n in zip(A, N):
        ra, m, k, ra2 = a.shape
        D = np.einsum("amkA,ij->amikjA", a, np.eye(n)).reshape(ra, m * n, k * n, ra2)
        out.append(D)
    return out


# --------------------------------------------------------------------------- DMRG matvec
_U64 = np.uint64


def _splitmix64(x):
    """SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        x = x + _U64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> _U64(30))) * _U64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> _U64(27))) * _U64(0x94D049BB133111EB)
    return x ^ (x >> _U64(31))


def randn_stream(seed, stream_id, n):
    """n normal numbers of counter stream `stream_id`: element i is Box-Muller of two SplitMix64
    words keyed by (seed, stream_id, i) -- the rule of the device's normal generator, restated
    so that the oracle draws the same kick columns as the GPU sweep."""
    with np.errstate(over="ignore"):
        i = np.arange(n, dtype=np.uint64)
        key = _U64(stream_id) * _U64(0x100000001B3) + i * _U64(2)
        a = _splitmix64(_U64(seed) ^ _splitmix64(key))
        b = _splitmix64(a + _U64(0x632BE59BD9B4E019))
    u1 = ((a >> _U64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740993.0)
    u2 = (b >> _U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def dmrg_mv(A, x, y0, nswp=20, eps=1e-12, rmax=32768, kickrank=4, seed=0, trace=None):
    """y ~= A x by DMRG sweeps (cpp/dmrg_mv.h:23-230, the reference's native fast_matvec).

    A: operator cores rA x M x N x rA'; x: cores rx x N x rx'; y0: initial cores (the reference
    draws rank-2 normal cores when none is given, :43-51). Each sweep: right-to-left QR
    orthogonalisation with the interfaces Phi (:87-108); then for every supercore (k, k+1) the
    projection W of A x (or y_k y_k+1 in the last sweep), delta = ||W - y_k y_k+1|| / ||W||
    driving the rank enlargement (:145-158), SVD truncated by rank_chop at ||W|| eps /
    d^(1.5, 0.5 in the last sweep) plus the enlargement (:160-172), kickrank normal columns
    appended by QR except in sweep nswp-1 (:177-187), Phi update (:200-209); the sweep after
    max delta < eps is the last (:212-220). The kick columns come from randn_stream(seed,
    (1 << 40) | (sweep << 20) | k), row-major m1 x kickrank -- the stream the GPU sweep draws
    them from. rmax: an int or a list of d+1 caps. `trace`, if a dict, receives per-sweep
    ranks and deltas."""
    d = len(x)
    M = [a.shape[1] for a in A]
    y = [np.array(c, dtype=np.float64, copy=True) for c in y0]
    ry = [1] + [c.shape[2] for c in y]
    caps = [1] + (d - 1) * [int(rmax)] + [1] if isinstance(rmax, int) else list(rmax)
    Phis = [None] * (d + 1)
    Phis[0] = np.ones((1, 1, 1))
    Phis[d] = np.ones((1, 1, 1))
    delta = [1.0] * (d - 1)
    delta_prev = [1.0] * (d - 1)
    r_enl = [2] * d
    last = False
    sweeps = 0
    if trace is not None:
        trace.setdefault("ranks", [])
        trace.setdefault("deltas", [])
    for swp in range(nswp):
        for k in range(d - 1, 0, -1):
            core = y[k].transpose(1, 2, 0).reshape(M[k] * ry[k + 1], ry[k])
            q, r = _qr(core)
            rnew = min(core.shape)
            y[k] = q.T.reshape(rnew, M[k], ry[k + 1])
            ry[k] = rnew
            y[k - 1] = (y[k - 1].reshape(-1, y[k - 1].shape[2]) @ r.T).reshape(-1, M[k - 1], rnew)
            phi = np.tensordot(Phis[k + 1], x[k], axes=([2], [2]))          # Y A' x n   (:99-101)
            phi = np.tensordot(A[k], phi, axes=([2, 3], [3, 1]))            # a m Y x
            Phis[k] = np.tensordot(y[k], phi, axes=([1, 2], [1, 2]))        # y a x
        for k in range(d - 1):
            W_prev = np.tensordot(y[k], y[k + 1], axes=([2], [0]))
            if not last:
                W1 = np.tensordot(Phis[k], x[k], axes=([2], [0]))              # y a n X   (:121-130)
                W1 = np.tensordot(A[k], W1, axes=([0, 2], [1, 2]))             # m A' y X
                W2 = np.tensordot(Phis[k + 2], x[k + 1], axes=([2], [2]))      # Y B X n'
                W2 = np.tensordot(A[k + 1], W2, axes=([2, 3], [3, 1]))         # A' m' Y X
                W = np.tensordot(W1, W2, axes=([1, 3], [0, 3])).transpose(1, 0, 2, 3)  # y m m' Y
            else:
                W = W_prev
            b = float(np.linalg.norm(W))
            if b > 0:
                delta[k] = float(np.linalg.norm(W - W_prev)) / b
            else:
                delta[k] = 0.0
            with np.errstate(divide="ignore", invalid="ignore"):
                ratio = np.float64(delta[k]) / np.float64(delta_prev[k])
            if ratio >= 1 and delta[k] > eps:
                r_enl[k] += 1
            if ratio < 0.1 and delta[k] < eps:
                r_enl[k] = max(r_enl[k] - 1, 1)
            m1 = W.shape[0] * W.shape[1]
            U, S, Vt = np.linalg.svd(W.reshape(m1, -1), full_matrices=False)
            r_new = rank_chop(S, b * eps / d ** (0.5 if last else 1.5))
            if not last:
                r_new += r_enl[k]
            r_new = max(min(r_new, S.shape[0], caps[k + 1]), 1)
            W1 = U[:, :r_new]
            W2 = Vt[:r_new, :].T * S[:r_new]
            if swp < nswp - 1:
                sid = (1 << 40) | (swp << 20) | k
                kick = randn_stream(seed, sid, m1 * kickrank).reshape(m1, kickrank)
                W1, Rm = _qr(np.concatenate([W1, kick], axis=1))
                W2 = np.concatenate([W2, np.zeros((W2.shape[0], kickrank))], axis=1)
                W2 = Rm @ W2.T
                r_new = W1.shape[1]
            else:
                W2 = W2.T
            ry[k + 1] = r_new
            y[k] = W1.reshape(ry[k], M[k], r_new)
            y[k + 1] = W2.reshape(r_new, M[k + 1], ry[k + 2])
            phi = np.tensordot(Phis[k], x[k], axes=([2], [0]))              # y a n X   (:200-203)
            phi = np.tensordot(phi, A[k], axes=([1, 2], [0, 2]))            # y X m A'
            Phis[k + 1] = np.tensordot(y[k], phi, axes=([0, 1], [0, 2])).transpose(0, 2, 1)  # Y A' X
        sweeps = swp + 1
        if trace is not None:
            trace["ranks"].append(list(ry))
            trace["deltas"].append(list(delta))
        if last:
            break
        if max(delta) < eps:
            last = True
        delta_prev = list(delta)
    if trace is not None:
        trace["sweeps"] = sweeps
    return y


# --------------------------------------------------------------------------- TT helpers
def tt_full(cores):
    """Dense tensor of a TT vector (cores r x n x r') -- reference cpp/full.h:3-12."""
    t = cores[0]
    for c in cores[1:]:
        t = np.tensordot(t, c, axes=([t.ndim - 1], [0]))
    return t.reshape(t.shape[1:-1])


def tt_matvec(A, x):
    """Cores of A @ x for a TT operator A and TT vector x (rank rA*rx)."""
    out = []
    for a, c in zip(A, x):
        ra, m, n, ra2 = a.shape
        rc, _, rc2 = c.shape
        w = np.einsum("amnb,cnd->acmbd", a, c)
        out.append(w.reshape(ra * rc, m, ra2 * rc2))
    return out


def tt_add(x, y, alpha=1.0, beta=1.0):
    """Cores of alpha*x + beta*y by block concatenation."""
    d = len(x)
    out = []
    for k in range(d):
        a, c = x[k], y[k]
        if k == 0:
            a = alpha * a
            c = beta * c
        ra, n, ra2 = a.shape
        rc, _, rc2 = c.shape
        if d == 1:
            out.append(a + c)
        elif k == 0:
            out.append(np.concatenate([a, c], axis=2))
        elif k == d - 1:
            out.append(np.concatenate([a, c], axis=0))
        else:
            z = np.zeros((ra + rc, n, ra2 + rc2))
            z[:ra, :, :ra2] = a
            z[ra:, :, ra2:] = c
            out.append(z)
    return out


def tt_norm(cores):
    """Frobenius norm by a left-to-right QR sweep (no squaring of small quantities)."""
    cur = cores[0]
    for k in range(len(cores) - 1):
        r0, n, r1 = cur.shape
        q, r = _qr(cur.reshape(r0 * n, r1))
        nxt = cores[k + 1]
        cur = (r @ nxt.reshape(nxt.shape[0], -1)).reshape(q.shape[1], nxt.shape[1], nxt.shape[2])
    return float(np.linalg.norm(cur))


def tt_residual(A, x, b):
    """||A x - b|| / ||b|| in TT arithmetic."""
    return tt_norm(tt_add(tt_matvec(A, x), b, 1.0, -1.0)) / tt_norm(b)

"""NumPy restatement of the reference CPU path (test oracle; see oracle/__init__.py).

Arrays use the reference's Julia shapes and column-major semantics: a 4D image
is (Nx, Ny, Nz, Nt); linear indices are Fortran-order (x fastest).

Precision modes of the POGM driver (`run_recon_arrays(..., mode=)`):
  "ref" : reference-faithful -- FP32 patches and SVD, POGM state promoted to
          complex128 by Julia's Float64 alpha (scripts/reconstruct.jl:159,
          src/mirt_mod.jl:112,131), i.e. the reference CPU path.
  "f32" : everything complex64 (the arithmetic model of the GPU drop-in).
  "f64" : everything complex128 (the FP64 oracle).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

AX3 = (0, 1, 2)
EPS32 = float(np.finfo(np.float32).eps)
EPS64 = float(np.finfo(np.float64).eps)

PINS = {
    "Asense_forward": "closed form vs numpy DFT (1 coil, S=1, full mask); brute-force DFT on tiny sizes",
    "Asense_adjoint": "dot test <Ax,y> = <x,A'y>; A'A = I for full mask + unit-RSS smaps",
    "img2patches/patches2img": "brute-force index loops; round trip; coverage-count closed form",
    "SVST": "closed forms: rank-1, 1xNt group shrink, orthogonal columns, beta=0, beta>=s1, zero",
    "patch_nucnorm": "sum of singular values vs closed forms",
    "pogm_mod": "scalar-recurrence golden table; toy quadratic with identity prox vs hand trace",
    "poweriter_mod": "diagonal operator with known sigma1; full-mask SENSE -> 1",
    "lambdas": "closed-form values for the configs",
    "svst_jvp": "Y equals the SVD route (SVST); tangents vs FP64 central differences; rank-1 closed "
                "forms (radial direction, beta direction); divided differences vs (g_i - g_j)/(l_i - l_j)",
    "pogm_lambda_jvp": "X equals run_recon_arrays(mode f64); dloss/dlambda vs FP64 central differences "
                       "of the same loss (restart sequence unchanged)",
    "tSNR": "constant series (std 0 -> mean/eps), two-level series with known mean and std, "
            "magnitude taken of complex values",
    "detrend": "a + b t -> a exactly; constant series unchanged; the residual is orthogonal to "
               "(t - mean t); pinv form equals the centred least-squares slope",
    "recon_costs": "X = 0 -> (||y||^2/2, 0); noise-free data at the truth -> dc 0; rank-1 single-patch "
                   "scale -> lambda |u||v|; dc quadratic in X; reg positively homogeneous",
}


# --------------------------------------------------------------------------- FFT helpers

def _workers():
    import os
    return max(1, min(16, os.cpu_count() or 1))


def cfft3(x, sign=-1):
    """Unnormalised 3D DFT over axes 0..2 with exponent sign `sign`."""
    if sign < 0:
        return sfft.fftn(x, axes=AX3, workers=_workers())
    return sfft.ifftn(x, axes=AX3, norm="forward", workers=_workers())  # unnormalised inverse


# --------------------------------------------------------------------------- SENSE
# src/sense_gpu.jl:56-108 (the GPU file's convention, which it states matches MIRT.Asense)

def sample_index(samp):
    """0-based linear indices of sampled voxels in column-major order (src/sense_gpu.jl:68)."""
    return np.flatnonzero(np.asarray(samp).ravel(order="F"))


def Asense_forward(x, samp, smaps, fft_forward=True, unitary=True):
    """y (K, Nc) = scale * fftshift(fft(ifftshift(smaps .* x)))[samp, :]  (src/sense_gpu.jl:71-81)."""
    N = smaps.shape[:3]
    Nc = smaps.shape[3]
    rdt = np.float64 if x.dtype == np.complex128 or smaps.dtype == np.complex128 else np.float32
    scale = rdt(1.0 / math.sqrt(np.prod(N))) if unitary else rdt(1.0)
    xc = smaps * x[..., None]
    sh = sfft.ifftshift(xc, axes=AX3)
    k = cfft3(sh, -1 if fft_forward else +1)
    if not fft_forward:
        k = k / np.prod(N)  # src/sense_gpu.jl:77 uses the normalised ifft
    kc = scale * sfft.fftshift(k, axes=AX3)
    idx = sample_index(samp)
    return kc.reshape(-1, Nc, order="F")[idx, :]


def Asense_adjoint(y, samp, smaps, fft_forward=True, unitary=True, exact=True):
    """x (N) = sum_c conj(S_c) .* scale * ifftshift(IDFT(fftshift(zero-fill(y)))).

    exact=True uses the unnormalised inverse DFT, which makes this the conjugate
    transpose of Asense_forward.  exact=False reproduces src/sense_gpu.jl:93 as
    written (normalised ifft, i.e. 1/prod(N) times the adjoint).
    """
    N = smaps.shape[:3]
    Nc = smaps.shape[3]
    V = int(np.prod(N))
    dt = np.result_type(y.dtype, smaps.dtype)
    rdt = np.float64 if dt == np.complex128 else np.float32
    scale = rdt(1.0 / math.sqrt(V)) if unitary else rdt(1.0)
    full = np.zeros((V, Nc), dtype=dt)
    full[sample_index(samp), :] = y
    full = full.reshape(*N, Nc, order="F")
    if exact:
        # (fftshift o F o ifftshift)^H = fftshift o F^H o ifftshift; for even sizes
        # this equals the reference's ifftshift o F^H o fftshift.
        xc = cfft3(sfft.ifftshift(full, axes=AX3), +1 if fft_forward else -1)
        if not fft_forward:
            xc = xc / V
        xc = scale * sfft.fftshift(xc, axes=AX3)
    else:
        # src/sense_gpu.jl:92-96 as written
        sh = sfft.fftshift(full, axes=AX3)
        xc = cfft3(sh, +1) / V if fft_forward else cfft3(sh, -1)
        xc = scale * sfft.ifftshift(xc, axes=AX3)
    return np.sum(np.conj(smaps) * xc, axis=3)


def A_forward(x4, mask4, smaps, **kw):
    """block_diag over frames (scripts/reconstruct.jl:142): (Nx,Ny,Nz,Nt) -> (K,Nc,Nt)."""
    Nt = x4.shape[3]
    return np.stack([Asense_forward(x4[..., t], mask4[..., t], smaps, **kw) for t in range(Nt)], axis=2)


def A_adjoint(y3, mask4, smaps, **kw):
    Nt = y3.shape[2]
    return np.stack([Asense_adjoint(y3[:, :, t], mask4[..., t], smaps, **kw) for t in range(Nt)], axis=3)


# --------------------------------------------------------------------------- patches
# src/recon.jl:53-126

def patch_geometry(N, patch_size, stride_size):
    ps = [min(int(patch_size[a]), int(N[a])) for a in range(3)]
    nsteps = [-(-(int(N[a]) - ps[a]) // int(stride_size[a])) for a in range(3)]  # cld
    Np = (nsteps[0] + 1) * (nsteps[1] + 1) * (nsteps[2] + 1)
    return ps, nsteps, Np


def _patch_starts(N, ps, ss, nsteps):
    """0-based starts in patch order (x fastest): min(i*s, N - ps)  (src/recon.jl:69-72)."""
    for iz in range(nsteps[2] + 1):
        for iy in range(nsteps[1] + 1):
            for ix in range(nsteps[0] + 1):
                yield (min(ix * ss[0], N[0] - ps[0]), min(iy * ss[1], N[1] - ps[1]),
                       min(iz * ss[2], N[2] - ps[2]))


def img2patches(img, patch_size, stride_size, dtype=np.complex64):
    """P (p, Nt, Np); src/recon.jl:53-78 (always ComplexF32 there: dtype=complex64)."""
    N = img.shape[:3]
    Nt = img.shape[3]
    ps, nsteps, Np = patch_geometry(N, patch_size, stride_size)
    p = ps[0] * ps[1] * ps[2]
    P = np.zeros((p, Nt, Np), dtype=dtype)
    for ip, (sx, sy, sz) in enumerate(_patch_starts(N, ps, stride_size, nsteps)):
        patch = img[sx:sx + ps[0], sy:sy + ps[1], sz:sz + ps[2], :]
        P[:, :, ip] = patch.reshape(p, Nt, order="F")
    return P


def patches2img(P, patch_size, stride_size, og_size, dtype=np.complex64):
    """Sum of patches in patch order / max(count, 1)  (src/recon.jl:96-126)."""
    Nt = P.shape[1]
    N = tuple(int(n) for n in og_size)
    ps, nsteps, Np = patch_geometry(N, patch_size, stride_size)
    rdt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    img = np.zeros((*N, Nt), dtype=dtype)
    cnt = np.zeros(N, dtype=rdt)
    for ip, (sx, sy, sz) in enumerate(_patch_starts(N, ps, stride_size, nsteps)):
        patch = P[:, :, ip].reshape(ps[0], ps[1], ps[2], Nt, order="F")
        img[sx:sx + ps[0], sy:sy + ps[1], sz:sz + ps[2], :] += patch
        cnt[sx:sx + ps[0], sy:sy + ps[1], sz:sz + ps[2]] += rdt(1)
    cnt = np.maximum(cnt, rdt(1))
    # Julia's ComplexF32 / Float32 divides the parts separately: Complex(re/x, im/x)
    # (NumPy's complex / real division rounds differently).
    img.real /= cnt[..., None]
    img.imag /= cnt[..., None]
    return img


def coverage_count(N, patch_size, stride_size):
    ps, nsteps, _ = patch_geometry(N, patch_size, stride_size)
    cnt = np.zeros(tuple(N), dtype=np.int64)
    for sx, sy, sz in _patch_starts(N, ps, stride_size, nsteps):
        cnt[sx:sx + ps[0], sy:sy + ps[1], sz:sz + ps[2]] += 1
    return cnt


# --------------------------------------------------------------------------- SVST
# src/recon.jl:183-269

def SVST(X, beta):
    """U * Diagonal(max(S - beta, 0)) * Vt with the zero shortcut (src/recon.jl:183-205)."""
    if np.linalg.norm(X) == 0:
        return np.zeros_like(X)
    U, S, Vh = np.linalg.svd(X, full_matrices=False)  # LAPACK gesdd, as Julia's svd
    bT = S.dtype.type(beta)
    s = np.maximum(S - bT, S.dtype.type(0))
    keep = s > 0
    if not keep.any():
        return np.zeros((U.shape[0], Vh.shape[1]), dtype=X.dtype)
    return ((U[:, keep] * s[keep]) @ Vh[keep, :]).astype(X.dtype, copy=False)


def svst_batched(P, beta):
    out = np.empty_like(P)
    b = np.broadcast_to(np.asarray(beta, dtype=np.float64), (P.shape[2],))
    for i in range(P.shape[2]):
        out[:, :, i] = SVST(np.array(P[:, :, i]), b[i])
    return out


def patchSVST(img, beta, patch_size, stride_size, dtype=np.complex64):
    """src/recon.jl:251-269 (scalar or per-patch beta)."""
    P = img2patches(img, patch_size, stride_size, dtype=dtype)
    P = svst_batched(P, beta)
    return patches2img(P, patch_size, stride_size, img.shape[:3], dtype=dtype)


def patch_nucnorm(P, lambdas=None):
    """src/recon.jl:138-166."""
    Np = P.shape[2]
    costs = np.zeros(Np)
    for i in range(Np):
        s = np.linalg.svd(P[:, :, i], compute_uv=False)
        if lambdas is None:
            costs[i] = float(np.sum(s, dtype=np.float64))
        else:
            costs[i] = lambdas[i] * float(np.sum(s, dtype=np.float64)) if s[0] > 0 else 0.0
    return float(np.sum(costs)) if lambdas is None else float(np.sum(costs)) / Np


# --------------------------------------------------------------------------- POGM
# src/mirt_mod.jl:26-29, 74-188

def _gr_restart(Fgrad, ynew_yold, restart_cutoff):
    """sum(Float64, real(-Fgrad .* ynew_yold)) <= cutoff*norm(Fgrad)*norm(ynew_yold)  (no conjugate)."""
    lhs = float(np.sum(np.real(-Fgrad * ynew_yold), dtype=np.float64))
    return lhs <= restart_cutoff * float(np.linalg.norm(Fgrad)) * float(np.linalg.norm(ynew_yold))


def pogm_mod(x0, Fcost, f_grad, f_L, *, f_mu=0.0, mom="pogm", restart="gr", restart_cutoff=0.0,
             bsig=1, niter=10, g_prox=lambda z, c: z, fun=lambda it, xk, yk, r: None,
             scalar=np.float64):
    """Statement-by-statement restatement of pogm_mod (src/mirt_mod.jl:74-188).

    `scalar` is the type of the step scalars: np.float64 reproduces Julia's
    promotion of ComplexF32 arrays to ComplexF64 (Float64 alpha); np.float32
    keeps complex64 state.  g_prox may mutate its argument (the reference's
    closure does), which makes xnew and znew the same array as in the reference.
    """
    if mom not in ("pgm", "fpgm", "pogm"):
        raise ValueError(f"mom={mom}")
    if restart not in ("none", "gr", "fr"):
        raise ValueError(f"restart={restart}")
    if not f_L >= 0:
        raise ValueError("f_L")
    if not f_mu >= 0:
        raise ValueError("f_mu")
    if not bsig >= 0:
        raise ValueError("bsig")
    if not abs(restart_cutoff) < 1:
        raise ValueError("restart_cutoff")
    S = scalar
    L = f_L
    mu = f_mu
    q = mu / L
    told = 1.0
    sig = 1.0
    zetaold = 1.0
    xold = x0
    yold = x0
    uold = x0
    zold = x0
    Fcostold = Fcost(x0)
    Fgradold = np.zeros(x0.shape)
    out = [None] * (niter + 1)
    out[0] = fun(0, x0, x0, False)
    xnew = x0
    ynew = x0
    tnew = told
    for it in range(1, niter + 1):
        alpha = 2 / (L + mu) if (mom == "pgm" and mu != 0) else 1 / L
        a = S(alpha)
        fgrad = f_grad(xold)
        is_restart = False
        if mom in ("pgm", "fpgm"):
            ynew = g_prox(xold - a * fgrad, alpha)
            Fgrad = -S(1 / alpha) * (ynew - xold)
            Fcostnew = Fcost(ynew)
            if restart != "none":
                if (restart == "fr" and Fcostnew > Fcostold) or \
                   (restart == "gr" and _gr_restart(Fgrad, ynew - yold, restart_cutoff)):
                    told = 1.0
                    is_restart = True
                Fcostold = Fcostnew
        else:
            unew = xold - a * fgrad
        if mom == "fpgm" and mu != 0:
            beta = (1 - math.sqrt(q)) / (1 + math.sqrt(q))
        elif mom == "pogm" and mu != 0:
            beta = (2 + q - math.sqrt(q ** 2 + 8 * q)) ** 2 / 4 / (1 - q)
        elif mom != "pgm":
            tnew = 0.5 * (1 + math.sqrt(1 + 8 * told ** 2)) if (mom == "pogm" and it == niter) \
                else 0.5 * (1 + math.sqrt(1 + 4 * told ** 2))
            beta = (told - 1) / tnew
        if mom == "pgm":
            xnew = ynew
        elif mom == "fpgm":
            xnew = ynew + S(beta) * (ynew - yold)
        else:
            gamma = (2 + q - math.sqrt(q ** 2 + 8 * q)) / 2 if mu != 0 else sig * told / tnew
            znew = (unew + S(beta) * (unew - uold) + S(gamma) * (unew - xold)
                    - S(beta * alpha / zetaold) * (xold - zold))
            zetanew = alpha * (1 + beta + gamma)
            xnew = g_prox(znew, zetanew)
            Fgrad = fgrad - S(1 / zetanew) * (xnew - znew)
            ynew = xold - a * Fgrad
            Fcostnew = Fcost(xnew)
            if restart != "none":
                if (restart == "fr" and Fcostnew > Fcostold) or \
                   (restart == "gr" and _gr_restart(Fgrad, ynew - yold, restart_cutoff)):
                    tnew = 1.0
                    sig = 1.0
                    is_restart = True
                elif float(np.sum(np.real(Fgrad * Fgradold), dtype=np.float64)) < 0:
                    sig = bsig * sig
                Fcostold = Fcostnew
                Fgradold = Fgrad
            uold = unew
            zold = znew
            zetaold = zetanew
        out[it] = fun(it, xnew, ynew, is_restart)
        xold = xnew
        yold = ynew
        if mom != "pgm" and mu == 0:
            told = tnew
    return (xnew if mom == "pogm" else ynew), out


def pogm_scalars(niter, restarts=()):
    """(t_k, beta_k, gamma_k, zeta/alpha) of the aliased POGM with the given restart iterations
    (src/mirt_mod.jl:140-156,166-168,184), for golden tables."""
    told, sig = 1.0, 1.0
    rows = []
    rs = set(restarts)
    for it in range(1, niter + 1):
        tnew = 0.5 * (1 + math.sqrt(1 + 8 * told ** 2)) if it == niter else 0.5 * (1 + math.sqrt(1 + 4 * told ** 2))
        beta = (told - 1) / tnew
        gamma = sig * told / tnew
        rows.append((tnew, beta, gamma, 1 + beta + gamma))
        if it in rs:
            tnew, sig = 1.0, 1.0
        told = tnew
    return rows


def poweriter_mod(Afwd, Aadj, n, *, niter=200, tol=1e-6, x0=None, dtype=np.complex128):
    """src/mirt_mod.jl:215-238 with A given as forward/adjoint callables."""
    x = np.ones(n, dtype=dtype) if x0 is None else np.array(x0, dtype=dtype)
    ratio_old = np.inf
    it_done = niter
    for it in range(1, niter + 1):
        Ax = Afwd(x)
        ratio = np.linalg.norm(Ax) / np.linalg.norm(x)
        if abs(ratio - ratio_old) / ratio < tol:
            it_done = it
            break
        ratio_old = ratio
        x = Aadj(Ax)
        x = x / np.linalg.norm(x)
    return x, float(np.linalg.norm(Afwd(x)) / np.linalg.norm(x)), it_done


# --------------------------------------------------------------------------- driver
# scripts/reconstruct.jl:86-226 (without file I/O)

def lambdas_ong_lustig(N, Nt, patch_sizes):
    """scripts/reconstruct.jl:163-169."""
    Nvox = int(np.prod(N))
    return [math.sqrt(int(np.prod(ps))) + math.sqrt(Nt) + math.sqrt(math.log(Nvox * Nt / max(int(np.prod(ps)), Nt)))
            for ps in patch_sizes]


def normalise_smaps(smaps_raw):
    """smaps ./ (sqrt.(sum(abs2.(smaps); dims=4)) .+ eps(Float32)) in Float32
    (scripts/reconstruct.jl:88-89): abs2 = re*re + im*im, summed over the coils in order
    (Julia's reduction over a trailing dimension), each part divided by the result."""
    s = np.asarray(smaps_raw).astype(np.complex64)
    acc = np.zeros(s.shape[:3], dtype=np.float32)
    for c in range(s.shape[3]):
        re, im = s[..., c].real, s[..., c].imag
        acc = acc + (re * re + im * im)
    den = (np.sqrt(acc) + np.float32(EPS32))[..., None]
    out = np.empty_like(s)
    out.real = s.real / den
    out.imag = s.imag / den
    return out


def sense_comb(ksp, smaps):
    """src/recon.jl:312-323: fftshift(ifft(ifftshift(ksp), 1:3)) combined with conj(smaps)/(RSS^2+eps)."""
    img_mc = sfft.fftshift(sfft.ifftn(sfft.ifftshift(ksp, axes=AX3), axes=AX3, workers=_workers()), axes=AX3)
    denom = np.sum(np.abs(smaps) ** 2, axis=3).astype(np.float64) + EPS64
    Nt = ksp.shape[4]
    img = np.zeros((*ksp.shape[:3], Nt), dtype=ksp.dtype)
    for t in range(Nt):
        num = np.sum(np.conj(smaps) * img_mc[..., t], axis=3)
        img[..., t] = num / denom
    return img


def quantile_linear(a, q):
    """Statistics.quantile (type 7) == numpy 'linear'."""
    return float(np.quantile(np.asarray(a, dtype=np.float64), q, method="linear"))


def prepare(ksp0, smaps_raw, N, Nvc, Nt):
    """Setup of scripts/reconstruct.jl:86-118,145-147: returns smaps, normalised ksp (K,Nvc,Nt), mask, R, sf."""
    smaps = normalise_smaps(smaps_raw)
    ksp0 = np.asarray(ksp0).astype(np.complex64)
    assert ksp0.shape == (*N, Nvc, Nt), f"Unexpected k-space shape: {ksp0.shape}"
    img0 = sense_comb(ksp0, smaps)
    sf = quantile_linear(np.abs(img0).ravel(order="F"), 0.99)
    # ksp0 ./= max(sf, eps(Float32)): ComplexF32 ./ Float64 divides in Float64, stores Float32
    ksp0 = (ksp0.astype(np.complex128) / max(sf, EPS32)).astype(np.complex64)
    omega = ksp0[:, :, :, 0, :] != 0
    R = float(np.prod(N)) / float(np.sum(omega[..., 0]))
    for ic in range(1, Nvc):
        assert np.array_equal(omega, ksp0[:, :, :, ic, :] != 0), f"Coil {ic + 1} has a different sampling pattern"
    for it in range(1, Nt):
        assert np.sum(omega[..., it]) == np.sum(omega[..., it - 1]), f"Frame {it + 1} has a different sample count"
    flat = ksp0.reshape(-1, Nvc, Nt, order="F")
    ksp = np.stack([flat[sample_index(omega[..., t]), :, t] for t in range(Nt)], axis=2)
    return smaps, ksp, omega, R, sf


def run_recon_arrays(ksp, omega, smaps, *, N, Nt, PATCH_SIZES, STRIDES, NITERS, sigma1A=1.0,
                     mode="ref", exact_adjoint=True, fun=None, lambdas=None, restart="gr",
                     restart_cutoff=0.0):
    """POGM part of run_recon (scripts/reconstruct.jl:159-226) on prepared inputs.
    lambdas (optional) replaces the Ong-Lustig weights (for derivative checks);
    restart / restart_cutoff are pogm_mod's options (run_recon keeps the defaults
    :gr and 0, src/mirt_mod.jl:81-82).

    Returns (X (Nx,Ny,Nz,Nt,S), restarts list, extras dict).
    """
    S = len(PATCH_SIZES)
    L = S * sigma1A ** 2
    lams = list(lambdas) if lambdas is not None else lambdas_ong_lustig(N, Nt, PATCH_SIZES)
    cdt = np.complex128 if mode == "f64" else np.complex64
    pdt = np.complex128 if mode == "f64" else np.complex64
    scalar = np.float32 if mode == "f32" else np.float64
    sm = smaps.astype(cdt)
    y = ksp.astype(cdt)

    def Af(x4):
        return A_forward(x4.astype(cdt, copy=False), omega, sm)

    def Aa(y3):
        return A_adjoint(y3, omega, sm, exact=exact_adjoint)

    def image_sum(X):
        s = X[..., 0].copy()
        for k in range(1, S):
            s = s + X[..., k]
        return s

    def dc_cost_grad(X):
        g = Aa(Af(image_sum(X)) - y)
        g = g.astype(X.dtype, copy=False) if mode != "ref" else g
        return np.repeat(g[..., None], S, axis=4)

    def g_prox(X, c):
        for k in range(S):
            X[..., k] = patchSVST(X[..., k], c * lams[k], PATCH_SIZES[k], STRIDES[k], dtype=pdt)
        return X

    X0 = np.zeros((*N, Nt, S), dtype=cdt)
    X0[..., 0] = Aa(y)
    X, out = pogm_mod(X0, lambda X: 0.0, dc_cost_grad, L, mom="pogm", niter=NITERS, g_prox=g_prox,
                      fun=fun or (lambda it, xk, yk, r: r), scalar=scalar, restart=restart,
                      restart_cutoff=restart_cutoff)
    restarts = [bool(o) for o in out[1:]] if fun is None else out
    return X, restarts, {"L": L, "lambdas": lams}


def _svst_nuc(X, beta):
    """SVST(X, beta) (the same statements as SVST above) and the nuclear norm of its
    output, sum(max(S - beta, 0)) in FP64 (src/recon.jl:138-146 applied to the SVST
    output, whose singular values are the thresholded ones)."""
    if np.linalg.norm(X) == 0:
        return np.zeros_like(X), 0.0
    U, S, Vh = np.linalg.svd(X, full_matrices=False)
    bT = S.dtype.type(beta)
    s = np.maximum(S - bT, S.dtype.type(0))
    keep = s > 0
    if not keep.any():
        return np.zeros((U.shape[0], Vh.shape[1]), dtype=X.dtype), 0.0
    return ((U[:, keep] * s[keep]) @ Vh[keep, :]).astype(X.dtype, copy=False), float(np.sum(s, dtype=np.float64))


def _disjoint(N, patch_size, stride_size):
    return int(coverage_count(N, patch_size, stride_size).max()) <= 1


def prox_scale_nuc(z, beta, patch_size, stride_size):
    """patchSVST of one scale (src/recon.jl:251-269) in z's dtype, returning (image,
    nuclear norm of the output's patches or None).  The nuclear norm is exact only
    when the patches are disjoint (then each output patch is the SVST output); for
    overlapping patches None is returned.  1-voxel patches use the 1 x Nt closed form
    of the SVD (the group shrink v max(|v| - beta, 0) / |v|, pinned by
    test_svst_vector_group_shrink); the whole-image patch is thresholded without a
    patch copy."""
    N = z.shape[:3]
    Nt = z.shape[3]
    ps, nsteps, Np = patch_geometry(N, patch_size, stride_size)
    p = ps[0] * ps[1] * ps[2]
    dt = z.dtype
    if p == 1:
        v = z.reshape(-1, Nt, order="F")
        nrm = np.sqrt(np.sum(np.abs(v) ** 2, axis=1))
        rdt = nrm.dtype.type
        sh = np.maximum(nrm - rdt(beta), rdt(0))
        f = np.where(nrm > 0, sh / np.where(nrm > 0, nrm, rdt(1)), rdt(0))
        out = (v * f[:, None]).astype(dt, copy=False).reshape(z.shape, order="F")
        return out, float(np.sum(sh, dtype=np.float64))
    if Np == 1 and tuple(ps) == tuple(N):
        Y, nuc = _svst_nuc(z.reshape(p, Nt, order="F"), beta)
        return Y.reshape(z.shape, order="F"), nuc
    P = img2patches(z, patch_size, stride_size, dtype=dt)
    nuc = 0.0
    for i in range(Np):
        P[:, :, i], c = _svst_nuc(np.array(P[:, :, i]), beta)
        nuc += c
    out = patches2img(P, patch_size, stride_size, N, dtype=dt)
    return out, (nuc if _disjoint(N, patch_size, stride_size) else None)


def run_recon_lean(ksp, omega, smaps, *, N, Nt, PATCH_SIZES, STRIDES, NITERS, sigma1A=1.0,
                   dtype=np.complex128, restart="gr", restart_cutoff=0.0, lambdas=None,
                   log_costs=True, callback=None):
    """The POGM run of run_recon (scripts/reconstruct.jl:159-226) with the in-place
    prox, evaluated with two 5D arrays (X and U_old) instead of pogm_mod's ~10, so
    the production shapes fit in host memory.

    With g_prox mutating its argument, pogm_mod (src/mirt_mod.jl:111-185) computes
      g = A'(A sum_s X_s - y) (the same for every scale, reconstruct.jl:180-183),
      u = x - alpha g,  z = u + beta (u - u_old) + gamma (u - x) - beta alpha/zeta_old * 0
      (x_old and z_old are the same array), restart iff
      sum(Float64, real(-g .* (u - u_old))) <= cutoff |g|_5D |u - u_old| (:26-29, :163;
      Fgrad = fgrad and y_new = u_new because x_new === z_new), x_s = patchSVST(z_s, zeta lambda_s);
    this function evaluates exactly that, scale by scale.  Pinned against
    run_recon_arrays / pogm_mod (the statement-by-statement restatement) on small
    problems: equal iterates and restarts.

    log_costs: the logger (scripts/reconstruct.jl:176-191,205) of every iterate:
    dc_cost(x_k) from the residual of iteration k + 1's gradient (one more A for the
    last iterate) and reg_cost(x_k) from the thresholded singular values of the prox
    for disjoint patches (else a patch_nucnorm pass).
    callback(it, X, restart) is called after each iteration (it = 0 for X0).
    Returns (X, restarts, extras) with extras["dc_costs"], ["reg_costs"] when logged and
    extras["restart_cos"]: the restart test's lhs / (|Fgrad| |y_new - y_old|) per
    iteration, i.e. its margin to the cutoff."""
    S = len(PATCH_SIZES)
    L = S * sigma1A ** 2
    lams = list(lambdas) if lambdas is not None else lambdas_ong_lustig(N, Nt, PATCH_SIZES)
    rdt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    sm = smaps.astype(dtype)
    y = ksp.astype(dtype)
    alpha = 1.0 / L
    a = rdt(alpha)
    X = np.zeros((*N, Nt, S), dtype=dtype)
    X[..., 0] = A_adjoint(y, omega, sm)
    U = X.copy()
    dcs, regs = [], []

    def reg_of(k, img):
        return lams[k] * patch_nucnorm(img2patches(img, PATCH_SIZES[k], STRIDES[k], dtype=dtype))

    if log_costs:
        regs.append(sum(reg_of(k, X[..., k]) for k in range(S) if np.any(X[..., k])))
    if callback:
        callback(0, X, False)
    told, sig, restarts, cosines = 1.0, 1.0, [], []
    for it in range(1, NITERS + 1):
        s = X[..., 0].copy()
        for k in range(1, S):
            s = s + X[..., k]
        r = A_forward(s, omega, sm) - y
        del s
        if log_costs:
            dcs.append(0.5 * float(np.vdot(r.ravel(), r.ravel()).real))
        g = A_adjoint(r, omega, sm)
        del r
        tnew = 0.5 * (1 + math.sqrt(1 + 8 * told ** 2)) if it == NITERS else 0.5 * (1 + math.sqrt(1 + 4 * told ** 2))
        beta = (told - 1) / tnew
        gamma = sig * told / tnew
        zeta = alpha * (1 + beta + gamma)
        lhs = 0.0
        dnorm2 = 0.0
        reg = 0.0
        for k in range(S):
            u = X[..., k] - a * g
            du = u - U[..., k]
            lhs += float(np.sum(np.real(-g * du), dtype=np.float64))
            dnorm2 += float(np.vdot(du.ravel(), du.ravel()).real)
            z = u + rdt(beta) * du + rdt(gamma) * (u - X[..., k])
            del du
            U[..., k] = u
            del u
            xk, nuc = prox_scale_nuc(z, zeta * lams[k], PATCH_SIZES[k], STRIDES[k])
            del z
            X[..., k] = xk
            if log_costs:
                reg += lams[k] * (nuc if nuc is not None else patch_nucnorm(
                    img2patches(xk, PATCH_SIZES[k], STRIDES[k], dtype=dtype)))
            del xk
        is_restart = False
        gnorm = math.sqrt(S) * float(np.linalg.norm(g))
        cosines.append(lhs / (gnorm * math.sqrt(dnorm2)) if gnorm * dnorm2 > 0 else 0.0)
        if restart != "none":
            is_restart = lhs <= restart_cutoff * gnorm * math.sqrt(dnorm2)
        del g
        if is_restart:
            tnew, sig = 1.0, 1.0
        restarts.append(bool(is_restart))
        if log_costs:
            regs.append(reg)
        if callback:
            callback(it, X, is_restart)
        told = tnew
    extras = {"L": L, "lambdas": lams, "restart_cos": cosines}
    if log_costs:
        s = X.sum(axis=4)
        r = A_forward(s, omega, sm) - y
        dcs.append(0.5 * float(np.vdot(r.ravel(), r.ravel()).real))
        extras["dc_costs"] = dcs
        extras["reg_costs"] = regs
    return X, restarts, extras


def recon_costs(X, y, omega, smaps, lambdas, PATCH_SIZES, STRIDES, dtype=np.complex128):
    """run_recon's logger on one iterate X (Nx,Ny,Nz,Nt,S), scripts/reconstruct.jl:174-191,205:
    dc_cost = 0.5 ||A image_sum(X) - y||^2,
    reg_cost = sum_k lambda_k patch_nucnorm(img2patches(X[..., k], PATCH_SIZES[k], STRIDES[k])).
    Evaluated in `dtype` throughout (the reference: FP32 arrays, Float64 sums)."""
    S = X.shape[4]
    img = X[..., 0].astype(dtype)
    for k in range(1, S):
        img = img + X[..., k].astype(dtype)
    r = A_forward(img, omega, smaps.astype(dtype)) - y.astype(dtype)
    dc = 0.5 * float(np.vdot(r.ravel(), r.ravel()).real)
    reg = 0.0
    for k in range(S):
        P = img2patches(X[..., k].astype(dtype), PATCH_SIZES[k], STRIDES[k], dtype=dtype)
        reg += float(lambdas[k]) * patch_nucnorm(P)
    return dc, reg


# --------------------------------------------------------------------------- derivatives
# The reference has no autodiff; these are the FP64 forward-mode references of the
# differentiation form: directional derivatives of SVST (src/recon.jl:183-205) and
# of the unrolled aliased POGM (src/mirt_mod.jl:131-184 with the in-place prox of
# scripts/reconstruct.jl:196-202) with respect to the per-scale lambda_k.

def svst_divdiff(lam, beta):
    """Divided differences D_ij = (g(l_i) - g(l_j)) / (l_i - l_j) of the Gram spectral
    function g(l) = max(1 - beta / sqrt(l), 0) (D_ii = g'(l_i)), in forms without the
    difference in a denominator: both above beta^2: beta / (s_i s_j (s_i + s_j));
    one above: g(l_i) / (l_i - l_j) with l_i - l_j >= (s_i - beta)(s_i + beta); both
    below: 0.  A kink at l = beta^2 (the derivative there is one-sided)."""
    lam = np.maximum(np.asarray(lam, dtype=np.float64), 0.0)
    s = np.sqrt(lam)
    up = s > beta
    g = np.where(up, 1.0 - beta / np.where(up, s, 1.0), 0.0)
    si, sj = s[:, None], s[None, :]
    ui, uj = up[:, None], up[None, :]
    both = beta / np.where(ui & uj, si * sj * (si + sj), 1.0)
    dl = lam[:, None] - lam[None, :]
    one_i = g[:, None] / np.where(ui & ~uj, dl, 1.0)
    one_j = g[None, :] / np.where(~ui & uj, -dl, 1.0)
    return np.where(ui & uj, both, np.where(ui & ~uj, one_i, np.where(~ui & uj, one_j, 0.0)))


def svst_jvp(Z, beta, Zdot, bdot=0.0):
    """(Y, Ydot) for Y = SVST(Z, beta) = Z g(Z^H Z) and a direction (Zdot, bdot), FP64:
    G = V diag(l) V^H, W = V diag(g(l)) V^H, Y = Z W (= U diag(max(s - beta, 0)) V^H),
    Wdot = V (D o (V^H Gdot V) + diag(dg/dbeta) bdot) V^H with Gdot = Zdot^H Z + Z^H Zdot,
    Ydot = Zdot W + Z Wdot.  Uses the Gram of Z^H Z (p >= Nt) or of Z Z^H."""
    Z = np.asarray(Z, dtype=np.complex128)
    Zd = np.asarray(Zdot, dtype=np.complex128)
    left = Z.shape[0] < Z.shape[1]
    if left:  # Y^H = SVST(Z^H)
        Y, Yd = svst_jvp(Z.conj().T, beta, Zd.conj().T, bdot)
        return Y.conj().T, Yd.conj().T
    G = Z.conj().T @ Z
    lam, V = np.linalg.eigh(G)
    lam = np.maximum(lam, 0.0)
    s = np.sqrt(lam)
    up = s > beta
    g = np.where(up, 1.0 - beta / np.where(up, s, 1.0), 0.0)
    W = (V * g) @ V.conj().T
    Gd = Zd.conj().T @ Z + Z.conj().T @ Zd
    M = V.conj().T @ Gd @ V
    dgb = np.where(up, -1.0 / np.where(up, s, 1.0), 0.0)
    Wd = V @ (svst_divdiff(lam, beta) * M + np.diag(dgb * bdot)) @ V.conj().T
    return Z @ W, Zd @ W + Z @ Wd


def pogm_lambda_jvp(ksp, omega, smaps, x_ref, *, N, Nt, PATCH_SIZES, STRIDES, NITERS, lambdas=None,
                    sigma1A=1.0, progress=None):
    """Loss = mean |sum_s X_s - x_ref|^2 after NITERS iterations of the aliased POGM
    (the iteration of run_recon_arrays in mode f64: g = A'(A sum X - y), u = x - alpha g,
    z = u + beta (u - u_old) + gamma (u - x), x_s = patchSVST(z_s, zeta lambda_s);
    src/mirt_mod.jl:131-184, scripts/reconstruct.jl:174-226) and its derivative with
    respect to each lambda_k by forward-mode (one tangent per scale), FP64.
    The restart decisions are those of the primal run (piecewise constant in lambda).
    Returns (loss, [dloss/dlambda_k], restarts, X)."""
    S = len(PATCH_SIZES)
    L = S * sigma1A ** 2
    lams = list(lambdas) if lambdas is not None else lambdas_ong_lustig(N, Nt, PATCH_SIZES)
    sm = smaps.astype(np.complex128)
    y = ksp.astype(np.complex128)

    def AtA(x4):
        return A_adjoint(A_forward(x4, omega, sm), omega, sm)

    def prox_jvp(z, zd_list, thr, s):
        """patchSVST of one scale and its tangents (the patch maps are linear)."""
        ps, ss = PATCH_SIZES[s], STRIDES[s]
        P = img2patches(z, ps, ss, dtype=np.complex128)
        Pd = [img2patches(zd, ps, ss, dtype=np.complex128) for zd in zd_list]
        Y = np.empty_like(P)
        Yd = [np.empty_like(P) for _ in zd_list]
        for ip in range(P.shape[2]):
            for j in range(len(zd_list)):
                Yi, Ydj = svst_jvp(P[:, :, ip], thr, Pd[j][:, :, ip], zeta_now * (1.0 if j == s else 0.0))
                Yd[j][:, :, ip] = Ydj
            Y[:, :, ip] = Yi
        back = lambda Q: patches2img(Q, ps, ss, z.shape[:3], dtype=np.complex128)
        return back(Y), [back(Q) for Q in Yd]

    alpha = 1.0 / L
    X = np.zeros((*N, Nt, S), dtype=np.complex128)
    X[..., 0] = A_adjoint(y, omega, sm)
    Xd = [np.zeros_like(X) for _ in range(S)]  # d X / d lambda_j
    U, Ud = X.copy(), [np.zeros_like(X) for _ in range(S)]
    told, sig = 1.0, 1.0
    restarts = []
    for it in range(1, NITERS + 1):
        g = A_adjoint(A_forward(X.sum(axis=4), omega, sm) - y, omega, sm)
        gd = [AtA(Xd[j].sum(axis=4)) for j in range(S)]
        Un = X - alpha * g[..., None]
        Und = [Xd[j] - alpha * gd[j][..., None] for j in range(S)]
        tnew = 0.5 * (1 + math.sqrt(1 + 8 * told ** 2)) if it == NITERS else 0.5 * (1 + math.sqrt(1 + 4 * told ** 2))
        beta = (told - 1) / tnew
        gamma = sig * told / tnew
        zeta_now = alpha * (1 + beta + gamma)
        Z = Un + beta * (Un - U) + gamma * (Un - X)
        Zd = [Und[j] + beta * (Und[j] - Ud[j]) + gamma * (Und[j] - Xd[j]) for j in range(S)]
        rs = float(np.sum(np.real(-g[..., None] * (Un - U)), dtype=np.float64)) <= 0.0
        restarts.append(bool(rs))
        if rs:
            tnew, sig = 1.0, 1.0
        Xn = np.empty_like(X)
        Xnd = [np.empty_like(X) for _ in range(S)]
        for s in range(S):
            xs, xds = prox_jvp(Z[..., s], [Zd[j][..., s] for j in range(S)], zeta_now * lams[s], s)
            Xn[..., s] = xs
            for j in range(S):
                Xnd[j][..., s] = xds[j]
        X, Xd, U, Ud = Xn, Xnd, Un, Und
        told = tnew
        if progress:
            progress(it)
    r = X.sum(axis=4) - x_ref
    n = r.size
    loss = float(np.vdot(r.ravel(), r.ravel()).real) / n
    grads = [2.0 * float(np.vdot(r.ravel(), Xd[j].sum(axis=4).ravel()).real) / n for j in range(S)]
    return loss, grads, restarts, X


# --------------------------------------------------------------------------- analysis
# src/analysis.jl

def tSNR(img):
    """mean(|img|) / (std(|img|) + eps(Float32)) along the last axis, std with n - 1
    (src/analysis.jl:39-44).  |img| in Float32 as Julia's abs.(ComplexF32); the sums
    in FP64 (Julia's Float32 pairwise sums agree to ~1e-7)."""
    mag = np.abs(np.asarray(img).astype(np.complex64)).astype(np.float64)
    mean = mag.mean(axis=-1)
    sd = mag.std(axis=-1, ddof=1) if mag.shape[-1] > 1 else np.full(mean.shape, np.nan)
    return (mean / (sd + EPS32)).astype(np.float32)


def detrend(img):
    """detrend!(img) (src/analysis.jl:122-132) on a copy: per voxel b = (pinv(M) x)[2]
    with M = [ones(Nt) 1:Nt], then x -= (1:Nt) * b, in FP64, stored as ComplexF32."""
    x = np.asarray(img).astype(np.complex128)
    Nt = x.shape[-1]
    M = np.stack([np.ones(Nt), np.arange(1, Nt + 1, dtype=np.float64)], axis=1)
    P = np.linalg.pinv(M)
    b = x.reshape(-1, Nt) @ P[1]
    out = x.reshape(-1, Nt) - np.outer(b, M[:, 1])
    return out.reshape(x.shape).astype(np.complex64)


def nrmse(x, ref):
    ref = np.asarray(ref)
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))

"""Seeded inputs generated on the device (torch), for slabs too large for the
numpy generators of inputs.py.  Each returns X^T as an n x m row-major FP64
tensor, i.e. the column-major m x n slab with leading dimension m."""


def svd_input(torch, m, n, kappa, seed):
    """X = U diag(sigma) V^T: U (m x n) orthonormal (CholQR2 of a Gaussian),
    V (n x n) the Q factor of a Gaussian, sigma log-spaced from 1 to 1/kappa."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    G = torch.randn((n, m), generator=g, dtype=torch.float64, device="cuda")
    for _ in range(2):  # CholQR2: rows of G become orthonormal (U^T)
        L = torch.linalg.cholesky(G @ G.T)
        G = torch.linalg.inv(L) @ G
    V, _ = torch.linalg.qr(torch.randn((n, n), generator=g, dtype=torch.float64,
                                       device="cuda"))
    sig = torch.logspace(0, -float(torch.log10(torch.tensor(kappa))), n,
                         dtype=torch.float64, device="cuda")
    X = (V * sig) @ G
    del G
    return X.contiguous()


def glued_input(torch, m, n, s, delta, seed):
    """Blocks X_k = Q0 C_k + delta E_k with one fixed orthonormal m x s basis
    Q0 (CholQR2 of a Gaussian), C_k s x s and E_k m x s standard Gaussian."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    Q0 = torch.randn((s, m), generator=g, dtype=torch.float64, device="cuda")
    for _ in range(2):
        L = torch.linalg.cholesky(Q0 @ Q0.T)
        Q0 = torch.linalg.inv(L) @ Q0
    X = torch.empty((n, m), dtype=torch.float64, device="cuda")
    for k in range(n // s):
        C = torch.randn((s, s), generator=g, dtype=torch.float64, device="cuda")
        X[k * s:(k + 1) * s] = C.T @ Q0
        X[k * s:(k + 1) * s] += delta * torch.randn((s, m), generator=g, dtype=torch.float64,
                                                     device="cuda")
    del Q0
    return X


def to_host(X):
    """Device n x m row-major tensor -> numpy m x n column-major array."""
    return X.cpu().numpy().T

// Equilibrium density matrix by the contour sum of Zandpack/cmdtools/SCF:209-215,
//   ρ = Σ_i (i/2π) (W_i G_i - conj(W_i) G_i†),   G_i = (E_i - H - Σ(E_i))⁻¹,
// with the self-energy Σ(E_i) supported on the n_c orbitals corb.  Instead of one
// n x n inversion per contour point, H = V diag(e) V† (eigenpairs from the caller)
// and the Woodbury identity on corb give
//   G_i = V D_i V† + V T'_i K_i T_i V†,   D_i = diag(1/(E_i - e)),
//   T_i = V_c D_i (n_c x n),  T'_i = D_i V_c† (n x n_c),
//   K_i = (I - Σ_i G0_i,cc)⁻¹ Σ_i,  G0_i,cc = T_i V_c†,
// so that X = Σ_i W_i G_i = V (diag(Σ_i W_i D_i) + Σ_i W_i T'_i K_i T_i) V† is a
// handful of MFMA GEMMs (the big one n x (nE n_c) x n) and ρ = (i/2π)(X - X†).
#pragma once

namespace zh {

// T'[o, i nc + c] = d_i(e_o) conj(V[corb_c, o]),  T[i nc + c, o] = V[corb_c, o] d_i(e_o),
// d_i(e) = 1 / (E_i - e); one k block.  V (n, n) row-major, column o = eigenvector o.
__global__ void k_scf_factors(int n, int nE, int nc, const double2* E, const double* e,
                              const double2* V, const int* corb, double2* Tp, double2* T) {
  const long per = (long)nE * nc, total = (long)n * per;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int o = (int)(idx / per);
    const long ic = idx - (long)o * per;
    const int i = (int)(ic / nc), c = (int)(ic - (long)i * nc);
    const double a = E[i].x - e[o], b = E[i].y;
    const double den = a * a + b * b;
    const double2 d = make_double2(a / den, -b / den);
    const double2 v = V[(long)corb[c] * n + o];
    Tp[(long)o * per + ic] = cmul(d, make_double2(v.x, -v.y));
    T[ic * n + o] = cmul(v, d);
  }
}

// w[o] = Σ_i W_i / (E_i - e_o)
__global__ void k_scf_diag(int n, int nE, const double2* E, const double2* W, const double* e,
                           double2* w) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  double sr = 0.0, si = 0.0;
  for (int i = 0; i < nE; ++i) {
    const double a = E[i].x - e[o], b = E[i].y;
    const double den = a * a + b * b;
    const double2 t = cmul(W[i], make_double2(a / den, -b / den));
    sr += t.x;
    si += t.y;
  }
  w[o] = make_double2(sr, si);
}

// Vh = V† and Vc† (n, nc): Vch[o, c] = conj(V[corb_c, o])
__global__ void k_scf_adjoint(int n, int nc, const double2* V, const int* corb, double2* Vh,
                              double2* Vch) {
  const long total = (long)n * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long r = idx / n, c = idx - r * n;
    const double2 v = V[c * n + r];
    Vh[idx] = make_double2(v.x, -v.y);
    if (c < nc) {
      const double2 u = V[(long)corb[c] * n + r];
      Vch[r * nc + c] = make_double2(u.x, -u.y);
    }
  }
}

// out[z] = I - acc   (the Woodbury matrix I - Σ G0_cc)
struct EpiIdMinus {
  double2* out; int ld; long sC;
  __device__ inline void operator()(int z, int row, int col, double re, double im) const {
    out[z * sC + (long)row * ld + col] = make_double2((row == col ? 1.0 : 0.0) - re, -im);
  }
};

// out[z] = W[z] acc
struct EpiScaleZ {
  double2* out; long ldc; long sC; const double2* W;
  __device__ inline void operator()(int z, int row, int col, double re, double im) const {
    out[z * sC + (long)row * ldc + col] = cmul(W[z], make_double2(re, im));
  }
};

// out = acc + diag(w)
struct EpiAddDiag {
  double2* out; int n; const double2* w;
  __device__ inline void operator()(int, int row, int col, double re, double im) const {
    if (row == col) {
      re += w[row].x;
      im += w[row].y;
    }
    out[(long)row * n + col] = make_double2(re, im);
  }
};

// out += acc (the later contour chunks of Y)
struct EpiAddTo {
  double2* out; int n;
  __device__ inline void operator()(int, int row, int col, double re, double im) const {
    double2& o = out[(long)row * n + col];
    o = make_double2(o.x + re, o.y + im);
  }
};

// Solves M X = R for a batch of nc x nc systems (one workgroup each) by Gauss-Jordan
// elimination with partial pivoting in global memory: R <- X, M destroyed.
constexpr int kGJMax = 1024;
__global__ __launch_bounds__(256) void k_gauss_jordan(int nc, double2* Mb, double2* Rb) {
  __shared__ double sv[256];
  __shared__ int si[256];
  __shared__ double2 f[kGJMax];
  const long off = (long)blockIdx.x * nc * nc;
  double2* A = Mb + off;
  double2* B = Rb + off;
  const int tid = threadIdx.x, bd = blockDim.x;
  for (int j = 0; j < nc; ++j) {
    double best = -1.0;
    int bi = j;
    for (int r = j + tid; r < nc; r += bd) {
      const double2 a = A[(long)r * nc + j];
      const double m = a.x * a.x + a.y * a.y;
      if (m > best) { best = m; bi = r; }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int s = bd / 2; s > 0; s >>= 1) {
      if (tid < s) {
        const bool take = sv[tid + s] > sv[tid] || (sv[tid + s] == sv[tid] && si[tid + s] < si[tid]);
        if (take) { sv[tid] = sv[tid + s]; si[tid] = si[tid + s]; }
      }
      __syncthreads();
    }
    const int p = si[0];
    if (p != j) {
      for (int c = tid; c < nc; c += bd) {
        const double2 t = A[(long)j * nc + c];
        A[(long)j * nc + c] = A[(long)p * nc + c];
        A[(long)p * nc + c] = t;
        const double2 u = B[(long)j * nc + c];
        B[(long)j * nc + c] = B[(long)p * nc + c];
        B[(long)p * nc + c] = u;
      }
    }
    __syncthreads();
    const double2 piv = A[(long)j * nc + j];
    const double pd = piv.x * piv.x + piv.y * piv.y;
    const double2 inv = make_double2(piv.x / pd, -piv.y / pd);
    __syncthreads();
    for (int c = tid; c < nc; c += bd) {
      A[(long)j * nc + c] = cmul(A[(long)j * nc + c], inv);
      B[(long)j * nc + c] = cmul(B[(long)j * nc + c], inv);
    }
    for (int r = tid; r < nc; r += bd) f[r] = A[(long)r * nc + j];
    __syncthreads();
    for (long q = tid; q < (long)nc * nc; q += bd) {
      const int r = (int)(q / nc), c = (int)(q - (long)r * nc);
      if (r == j) continue;
      const double2 fr = f[r];
      const double2 a = cmul(fr, A[(long)j * nc + c]);
      const double2 b = cmul(fr, B[(long)j * nc + c]);
      A[q].x -= a.x; A[q].y -= a.y;
      B[q].x -= b.x; B[q].y -= b.y;
    }
    __syncthreads();
  }
}

// R[z] = I (nc x nc) for every batch entry z
__global__ void k_identity_batch(int nb, int nc, double2* R) {
  const long cc = (long)nc * nc, total = (long)nb * cc;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long q = idx % cc;
    const long r = q / nc, c = q - r * nc;
    R[idx] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
}

// Ph = P† for P (rows, cols): Ph[c, r] = conj(P[r, c])
__global__ void k_conj_transpose(int rows, int cols, const double2* P, double2* Ph) {
  const long total = (long)rows * cols;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long r = idx / cols, c = idx - r * cols;
    const double2 v = P[idx];
    Ph[c * rows + r] = make_double2(v.x, -v.y);
  }
}

// ρ = (X + X†)/2 (the Hermitian part of a product that is Hermitian up to rounding)
__global__ void k_herm_part(int n, const double2* X, double2* rho) {
  const long total = (long)n * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long r = idx / n, c = idx - r * n;
    const double2 x = X[idx], y = X[c * n + r];
    rho[idx] = make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
  }
}

// ρ = (i/2π)(X - X†)
__global__ void k_scf_rho(int n, const double2* X, double2* rho) {
  const long total = (long)n * n;
  const double s = 1.0 / (2.0 * M_PI);
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long r = idx / n, c = idx - r * n;
    const double2 x = X[idx], y = X[c * n + r];
    // X - conj(Xᵀ), times i/2π
    const double dr = x.x - y.x, di = x.y + y.y;
    rho[idx] = make_double2(-s * di, s * dr);
  }
}

}  // namespace zh

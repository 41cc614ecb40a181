"""matflat_amd -- MI355X (gfx950) batched implementation of MatFlat's hot path.

* ``matflat_amd.hip``      batched device API over libmatflat_hip.so (C ABI in
                           include/matflat_hip.h)
* ``matflat_amd.matflat``  MatFlat's own signatures (Factorization.Lu,
                           Factorization.Cholesky, Blas.SolveTriangular,
                           Blas.MulMatMat) on device buffers
* ``matflat_amd.autograd`` torch.autograd.Function for the SPD solve (used by
                           the gradcheck harness)

The compute path is the HIP library only; importing the submodules that call
it fails if the library has not been built.
"""
__version__ = "0.1.0"

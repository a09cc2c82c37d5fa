"""fem_taichi_amd -- MI355X-native hex8 FEM hot path behind the reference's API.

Drop-in classes (same names and return types as tianshao1992/FEM_taichi):

* ``Physics``  -- mesh + node-pair pattern (``Read`` attributes, GPU pre_process)
* ``Model``    -- ``Get_MatrixKL``/``Get_MatrixK``/``Get_MatrixM``/``Get_VectorF``
* ``Post``     -- ``Get_Stress``
* ``Static_solver`` -- GPU PCG forms in place of Pardiso
* ``Eigen_solver`` / ``Block_lancozs`` -- shift-invert block Lanczos on the GPU
* ``read_model`` / ``read_cdb`` / ``read_lis`` -- the reference's mesh and
  boundary file readers (``Read``'s inputs)

All compute runs in the hand-written HIP kernels of ``libfem355.so``
(include/fem355.h); PyTorch only provides device memory and streams.
"""
from .physics import Physics
from .model import Model
from .post import Post
from .solve_static import Static_solver
from .solve_eigen import Block_lancozs, Eigen_solver
from .read_cdb import read_cdb, read_lis, read_model

__all__ = ["Physics", "Model", "Post", "Static_solver", "Eigen_solver", "Block_lancozs",
           "read_cdb", "read_lis", "read_model"]

"""The package path GSpaRC's renderer imports the rasterizer from
(renderer.py:5: `from cuda_kernel_rasterize.cuda_rasterizer import
GaussianRasterizerSetting, GaussianRasterizer`).  Here that path is this
build's `cuda_rasterizer` package, under both names."""
import sys

import cuda_rasterizer

sys.modules[__name__ + ".cuda_rasterizer"] = cuda_rasterizer

"""pypty_amd — MI355X-native hot path of iterative multislice ptychography.

Drop-in mirrors of PyPty's hot-path modules, computing in hand-written
gfx950 HIP kernels through a C ABI (``libpypty_amd.so``):

  pypty_amd.fft              fft2 / ifft2 / shift_fft2 / ifft2_ishift
  pypty_amd.multislice_core  multislice / multislice_grads / scatteradd
  pypty_amd.objective        loss_and_direction (+ the PtychoPass engine)
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401


def library_path() -> str:
    return _lib.LIB_PATH

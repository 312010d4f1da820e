"""C-ABI boundary: ctypes loader and bindings for libmkgpu."""

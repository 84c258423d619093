"""HIP/gfx950 backend of oncolytica_amd: ctypes binding of liboncohip.so,
simulation handles, batches, the Python-AST -> HIP rule translator and the
HipBackend that implements the reference's backend protocol."""

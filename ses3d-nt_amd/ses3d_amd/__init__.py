"""Host-side Python interface of the MI355X Ses3d-NT hot path.

api    -- ctypes binding of include/ses3d_hip.h (libses3d_hip.so)
synth  -- seeded synthetic meshes, Earth models, sources and receivers
build  -- compiles the HIP library (and the Fortran host program)
"""

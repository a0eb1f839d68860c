"""Python front end of the constant-Q PSTD engine for MI355X.

``cqpy.host``   ctypes bindings of libcq_host.so (formats, builders, revolve)
``cqpy.engine`` ctypes bindings of libcq_hip.so (the HIP engine's C ABI)
``cqpy.synth``  seeded synthetic models and shot jobs
"""

"""Python side of the MI355X RTE+RRTMGP column-radiation library.

  synthetic  seeded workloads (k-distributions, cloud LUTs, profiles)
  api        binding of the class API through the C facade (include/rrtmgp_api.h)
  device     device arrays over the C ABI (include/rrtmgp_hip.h)
"""
from . import synthetic  # noqa: F401

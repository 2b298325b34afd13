"""kpp_amd — batched Rosenbrock integration of KPP chemical mechanisms on MI355X.

The mechanism compiler (csrc/kppgen) turns KPP mechanism files into HIP
device functions; libkpphip.so integrates many cells per launch behind the
C ABI of include/kpphip.h; ``solver`` is the Python binding of that ABI.
"""
from .solver import (  # noqa: F401
    PERSISTENT, PLAIN, RODAS3, RODAS4, ROS2, ROS3, ROS4, SUN_FIXED, SUN_TOD, KppError,
    Mechanism, check_params, default_ipar, library, mechanism, mechanisms, rosenbrock)

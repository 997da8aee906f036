"""Constants and the tracer-scheme validator used by the FCT path.

Values follow fesom_jax/config.py:37-41 (R_EARTH), :100 (DT_DEFAULT) and :132-168
(TracerConfig). Everything on this path is IEEE float64; indices are int32.
"""

from __future__ import annotations

from typing import NamedTuple

PI = 3.14159265358979          # FESOM's truncated pi (fesom_jax/config.py:37)
RAD = PI / 180.0
R_EARTH = 6367500.0            # fesom_jax/config.py:41
OMEGA = 2.0 * PI / 86400.0
CYCLIC_LENGTH_RAD = 2.0 * PI   # fesom_jax/config.py:55
DT_DEFAULT = 2400.0            # fesom_jax/config.py:100

# AB2 stabilisation (fesom_jax/tracer_adv.py:44-46)
EPS_AB = 0.1
AB_OLD = -(0.5 + EPS_AB)
AB_NEW = 1.5 + EPS_AB

# Zalesak limiter constants (fesom_jax/tracer_adv.py:192-193)
FCT_FLUX_EPS = 1e-16
FCT_BIG = 1e3

TRACER_HOR_SCHEME = "MFCT"
TRACER_VER_SCHEME = "QR4C"
TRACER_LIMITER = "FCT"
TRACER_NUM_ORD_HOR = 0
TRACER_NUM_ORD_VER = 1


class TracerConfig(NamedTuple):
    """Scheme selector with the same fields and validation as
    fesom_jax/config.py:144-168: only (MFCT, QR4C, FCT, 0, 1) is implemented."""

    hor_scheme: str = TRACER_HOR_SCHEME
    ver_scheme: str = TRACER_VER_SCHEME
    limiter: str = TRACER_LIMITER
    num_ord_hor: int = TRACER_NUM_ORD_HOR
    num_ord_ver: int = TRACER_NUM_ORD_VER

    def validate(self) -> None:
        impl = (TRACER_HOR_SCHEME, TRACER_VER_SCHEME, TRACER_LIMITER,
                TRACER_NUM_ORD_HOR, TRACER_NUM_ORD_VER)
        got = (self.hor_scheme, self.ver_scheme, self.limiter,
               self.num_ord_hor, self.num_ord_ver)
        if got != impl:
            raise NotImplementedError(
                f"tracer scheme {got} != the implemented {impl}: the MFCT/QR4C "
                "reconstruction order is fixed in the kernels")

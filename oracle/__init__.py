"""CPU oracle for the RBRP column-ID path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline), never as a stand-in for the HIP
library.  The product path (``robust_blockwise_random_pivoting_amd``) never
imports it and raises if its HIP library is missing.

Every function cites the reference file:line it restates; see the module
docstring of :mod:`oracle.rbrp_oracle` for the pins that check each one.
"""

from .rbrp_oracle import (  # noqa: F401
    uniform_stream,
    sample_without_replacement,
    robust_filter,
    symmetric_difference,
    rbrp_colid,
    interpolation_matrix,
    id_error,
    RBRPResult,
    pivot_margin,
    filter_margin,
    first_near_tie,
    margin_summary,
    tie_pivot,
    TIE_CDF,
    TIE_STOP,
)
from .sketch_oracle import (  # noqa: F401
    compute_interpolation_matrix_os,
    rand_cpqr_os,
    rand_lupp_os,
    rand_cpqr,
    rand_lupp,
    rand_adap_lupp,
    pivots_to_permutation,
)
from .cpqr_oracle import cpqr_id, partial_time_fraction  # noqa: F401

"""MI355X (gfx950) execution path for reffine scan / filter / aggregate /
select / hash-probe queries.

  reffine_amd.reffiner   the builder DSL (mirror of reffine's reffiner.h)
  reffine_amd.runtime    device tables, compile_op, Query (C ABI via ctypes)
  reffine_amd.build      builds the in-tree C ABI library for gfx950
"""
from .reffiner import *  # noqa: F401,F403
from .runtime import (DeviceTable, Query, ReffineError, compile_check, compile_op,  # noqa: F401
                      device_info, init, load_arrow_file)

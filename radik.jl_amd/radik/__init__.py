"""RadiK for MI355X: radix top-k on hand-written gfx950 kernels.

Mirrors the exports of RadiK.jl (src/RadiK.jl:20-23): topk, topk_ (topk!),
topk_radix_select_ (topk_radix_select!), RadiKWorkspace; plus the batched
Plan and the serving Server.
"""
from ._lib import RadiKError, load  # noqa: F401
from .api import RadiKWorkspace, topk, topk_, topk_radix_select_  # noqa: F401
from .batched import Plan, Server  # noqa: F401
from .device import DeviceArray, Event, Stream, device_info, lds_caps, synchronize  # noqa: F401

__all__ = ["topk", "topk_", "topk_radix_select_", "RadiKWorkspace", "Plan", "Server", "DeviceArray",
           "Stream", "Event", "RadiKError"]

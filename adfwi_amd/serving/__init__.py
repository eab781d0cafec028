"""Continuous-batching serving loop for modelling and gradient requests (one GPU)."""
from .requests import Request, build_problem, execute, execute_pack, footprint_bytes, make_stream, pack_key
from .scheduler import Scheduler
from .loop import Outcome, ServingLoop, StreamReport, run_alone

__all__ = ["Request", "build_problem", "execute", "execute_pack", "footprint_bytes", "make_stream", "pack_key",
           "Scheduler", "Outcome", "ServingLoop", "StreamReport", "run_alone"]

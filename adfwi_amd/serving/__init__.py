"""Continuous-batching serving loop for modelling and gradient requests (one GPU)."""
from .requests import Request, build_problem, execute, footprint_bytes, make_stream
from .scheduler import Scheduler
from .loop import Outcome, ServingLoop, StreamReport, run_alone

__all__ = ["Request", "build_problem", "execute", "footprint_bytes", "make_stream", "Scheduler", "Outcome",
           "ServingLoop", "StreamReport", "run_alone"]

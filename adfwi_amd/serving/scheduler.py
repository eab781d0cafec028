"""Admission policy of the serving loop (host logic only, no device access).

Requests wait in arrival order.  A request is admitted when a slot is free and its HBM
footprint fits into what the running requests leave of the budget.  A request that does not
fit may be passed by later, smaller ones, but only `max_bypass` times: after that nothing is
admitted past it until it fits, so large requests cannot starve.  A request whose footprint
exceeds the whole budget is rejected at submission.
"""
from collections import OrderedDict, deque
from dataclasses import dataclass
from typing import Deque, List, Tuple


@dataclass
class _Waiting:
    req: object
    nbytes: int
    bypassed: int = 0


class Scheduler:
    def __init__(self, budget_bytes: int, max_slots: int, max_bypass: int = 4):
        if budget_bytes <= 0 or max_slots <= 0:
            raise ValueError("budget and slot count must be positive")
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.max_bypass = int(max_bypass)
        self.waiting: Deque[_Waiting] = deque()
        self.running: "OrderedDict[int, int]" = OrderedDict()
        self.rejected: List[Tuple[object, str]] = []
        self.peak_bytes = 0

    @property
    def used(self) -> int:
        return sum(self.running.values())

    def submit(self, req, nbytes: int) -> bool:
        if nbytes > self.budget:
            self.rejected.append((req, f"footprint {nbytes / 2**30:.1f} GiB exceeds the budget "
                                       f"{self.budget / 2**30:.1f} GiB"))
            return False
        self.waiting.append(_Waiting(req, int(nbytes)))
        return True

    def admit(self) -> List[object]:
        """Requests to start now, in arrival order."""
        out = []
        free = self.budget - self.used
        slots = self.max_slots - len(self.running)
        kept: Deque[_Waiting] = deque()
        blocked = False
        for w in self.waiting:
            if not blocked and slots > 0 and w.nbytes <= free:
                self.running[w.req.rid] = w.nbytes
                free -= w.nbytes
                slots -= 1
                out.append(w.req)
                for e in kept:          # every earlier request still waiting has been passed once more
                    e.bypassed += 1
                    if e.bypassed >= self.max_bypass:
                        blocked = True  # nothing passes it any more: it is next as soon as it fits
            else:
                kept.append(w)
                if w.nbytes > free and w.bypassed >= self.max_bypass:
                    blocked = True      # nothing passes it any more: it is next as soon as it fits
        self.waiting = kept
        self.peak_bytes = max(self.peak_bytes, self.used)
        return out

    def finish(self, rid: int) -> None:
        self.running.pop(rid, None)

    def idle(self) -> bool:
        return not self.waiting and not self.running

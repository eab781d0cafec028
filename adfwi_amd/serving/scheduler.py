"""Admission policy of the serving loop (host logic only, no device access).

Requests wait in arrival order.  A request is admitted when its slots are free and its HBM
footprint fits into what the running requests leave of the budget.  A request takes one slot,
or `weight` slots when submitted so (the serving loop gives long requests two of its four, so
that at most two of them share the device while short requests still run four at a time).  A
request that does not fit may be passed by later, smaller ones, but only `max_bypass` times:
after that nothing is admitted past it until it fits, so large requests cannot starve.  A
request whose footprint exceeds the whole budget is rejected at submission.

Packing: with a `pack_key`, a request being admitted takes later waiting requests with the
same key along (up to `max_pack_shots` shots in all, each within the free budget): the pack
shares one slot and one sequence of launches, and every member keeps its own footprint in the
accounting until the pack finishes.
"""
from collections import OrderedDict, deque
from dataclasses import dataclass
from typing import Callable, Deque, Dict, List, Optional, Tuple


@dataclass
class _Waiting:
    req: object
    nbytes: int
    bypassed: int = 0
    weight: int = 1


class Scheduler:
    def __init__(self, budget_bytes: int, max_slots: int, max_bypass: int = 4,
                 pack_key: Optional[Callable[[object], object]] = None, max_pack_shots: int = 64):
        if budget_bytes <= 0 or max_slots <= 0:
            raise ValueError("budget and slot count must be positive")
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.max_bypass = int(max_bypass)
        self.pack_key = pack_key
        self.max_pack_shots = int(max_pack_shots)
        self.waiting: Deque[_Waiting] = deque()
        self.running: "OrderedDict[int, int]" = OrderedDict()
        self._pack_of: Dict[int, int] = {}      # running request -> id of its pack (its first member)
        self._weight_of: Dict[int, int] = {}    # running pack -> slots it takes
        self.rejected: List[Tuple[object, str]] = []
        self.peak_bytes = 0

    @property
    def used(self) -> int:
        return sum(self.running.values())

    @property
    def slots_used(self) -> int:
        return sum(self._weight_of[p] for p in set(self._pack_of.values()))

    def submit(self, req, nbytes: int, weight: int = 1) -> bool:
        if nbytes > self.budget:
            self.rejected.append((req, f"footprint {nbytes / 2**30:.1f} GiB exceeds the budget "
                                       f"{self.budget / 2**30:.1f} GiB"))
            return False
        if not 1 <= weight <= self.max_slots:
            raise ValueError(f"slot weight {weight} outside 1..{self.max_slots}")
        self.waiting.append(_Waiting(req, int(nbytes), weight=int(weight)))
        return True

    def admit_packs(self) -> List[List[object]]:
        """Packs to start now (lists of requests sharing launches), in arrival order of their
        first members."""
        out = []
        free = self.budget - self.used
        slots = self.max_slots - self.slots_used
        kept: Deque[_Waiting] = deque()
        blocked = False
        queue = list(self.waiting)
        taken = set()
        for i, w in enumerate(queue):
            if id(w) in taken:
                continue
            if not blocked and slots >= w.weight and w.nbytes <= free:
                pack = [w]
                free -= w.nbytes
                taken.add(id(w))
                key = self.pack_key(w.req) if self.pack_key is not None else None
                if key is not None:
                    shots = w.req.nshots
                    between: List[_Waiting] = []   # waiting requests after the lead, not taken
                    passed: List[_Waiting] = []    # those a member has been taken past
                    for e in queue[i + 1:]:
                        if id(e) in taken:
                            continue
                        if (e.nbytes <= free and shots + e.req.nshots <= self.max_pack_shots
                                and self.pack_key(e.req) == key):
                            if any(p.bypassed >= self.max_bypass for p in between):
                                break           # a request that may not be passed again is ahead
                            pack.append(e)
                            taken.add(id(e))
                            free -= e.nbytes
                            shots += e.req.nshots
                            passed = list(between)
                        else:
                            between.append(e)
                    for p in passed:            # taking a member past a request passes it once
                        p.bypassed += 1
                slots -= w.weight
                self._weight_of[w.req.rid] = w.weight
                for m in pack:
                    self.running[m.req.rid] = m.nbytes
                    self._pack_of[m.req.rid] = w.req.rid
                out.append([m.req for m in pack])
                for e in kept:          # every earlier request still waiting has been passed once more
                    e.bypassed += 1
                    if e.bypassed >= self.max_bypass:
                        blocked = True  # nothing passes it any more: it is next as soon as it fits
            else:
                kept.append(w)
                if (w.nbytes > free or w.weight > slots) and w.bypassed >= self.max_bypass:
                    blocked = True      # nothing passes it any more: it is next as soon as it fits
        self.waiting = kept
        self.peak_bytes = max(self.peak_bytes, self.used)
        return out

    def admit(self) -> List[object]:
        """Requests to start now, in arrival order of their packs."""
        return [r for pack in self.admit_packs() for r in pack]

    def finish(self, rid: int) -> None:
        self.running.pop(rid, None)
        pack = self._pack_of.pop(rid, None)
        if pack is not None and pack not in self._pack_of.values():
            self._weight_of.pop(pack, None)

    def idle(self) -> bool:
        return not self.waiting and not self.running

"""Ensembles of independent simulations advanced together on one GPU.

Each member is an ordinary pyrandaSim (its own grid, context and fields) built
with `device_scalars=True`, so a time step — RK4 stages, EOM evaluation, the
dt reductions and the driver's dt update — runs without a host round trip.
`pyrandaEnsemble.capture()` records one time step of every member into a
single HIP graph (torch.cuda.CUDAGraph over the library's streams), each
member on its own stream forked from and joined back to the capture stream,
so members of uneven sizes run side by side; `step()` replays it. Small
members (a 128^2 KH layer is 16K points per field) are launch-bound one by
one: the graph removes the Python dispatch of every kernel and lets the
hardware overlap the members' kernels.

State across replays: the graph reads the tensors that hold each member's
state when it was captured and writes new tensors from its private pool, so
the captured step ends by copying every variable's new value (the driver's
time/dt and the FP64 states that FP32 members keep for their conservatively
filtered variables included) back into those state tensors.

No counterpart in the reference (pyranda runs one simulation per MPI
communicator, pyranda/pyranda.py:32-100).
"""
import torch


class pyrandaEnsemble:

    def __init__(self, members, advance, state):
        """members: pyrandaSim objects (device_scalars=True). advance(sim, st)
        runs one time step of `sim` and stores the driver's new scalars in
        the dict `st` (0-d device tensors, e.g. time and dt); state: one such
        dict per member, its initial values."""
        if len(members) != len(state):
            raise ValueError("one state dict per member")
        for m in members:
            if not m.PyMPI.device_scalars:
                raise ValueError("ensemble members need device_scalars=True")
        self.members = list(members)
        self.advance = advance
        self.state = [dict(st) for st in state]
        self.graph = None
        self._statics = None
        self.steps = 0

    # -- eager stepping -----------------------------------------------------
    def step_eager(self, nsteps=1):
        """Member by member, launch by launch (after capture() the results
        go back into the state tensors the graph reads, so eager steps and
        replays can be mixed)."""
        for _ in range(nsteps):
            for i, (sim, st) in enumerate(zip(self.members, self.state)):
                self.advance(sim, st)
                if self._statics is not None:
                    self._write_back(sim, st, self._statics[i], rebind=True)
            self.steps += 1

    @staticmethod
    def _write_back(sim, st, static, rebind):
        names, svars, sst, s64 = static
        for n in names:
            d = sim.variables[n].data
            if d is not svars[n]:
                svars[n].copy_(d)
                if rebind:
                    sim.variables[n].data = svars[n]
        for U, t in s64.items():
            d = sim._state64.get(U)
            if d is not None and d is not t:
                t.copy_(d)
            if rebind:
                sim._state64[U] = t
                sim._mark_state(U)
        for k, v in list(st.items()):
            if v is not sst[k]:
                sst[k].copy_(v)
                if rebind:
                    st[k] = sst[k]

    # -- graph capture --------------------------------------------------------
    def _static_state(self, sim, st):
        """Give every tensor-valued variable and driver scalar a tensor of its
        own (aliases are split) that the graph will read and write back."""
        seen = set()
        names = []
        # FP64 states of conserved variables (FP32 runs through the
        # conservative filter) that still match their fields
        fresh = [U for U in sim._state64 if sim._state_current(U)]
        for name, var in sim.variables.items():
            d = var.data
            if name.startswith("mesh") or not torch.is_tensor(d):
                continue
            if id(d) in seen:
                var.data = d.clone()
            seen.add(id(var.data))
            names.append(name)
        for k, v in st.items():
            if not torch.is_tensor(v) or id(v) in seen:
                st[k] = torch.as_tensor(v, dtype=torch.float64, device=sim.PyMPI.device).clone()
            seen.add(id(st[k]))
        s64 = {}
        for U in fresh:
            t = sim._state64[U]
            s64[U] = t.clone() if id(t) in seen else t
            seen.add(id(s64[U]))
            sim._state64[U] = s64[U]
            sim._mark_state(U)
        return names, {n: sim.variables[n].data for n in names}, dict(st), s64

    def capture(self, concurrent=True):
        """Record one time step of every member as one HIP graph. The members
        must have been stepped eagerly at least once (kernels compiled, work
        buffers allocated)."""
        dev = self.members[0].PyMPI.device
        torch.cuda.synchronize(dev)
        statics = [self._static_state(sim, st) for sim, st in zip(self.members, self.state)]
        streams = [torch.cuda.Stream(device=dev) for _ in self.members] if concurrent else None
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cap = torch.cuda.current_stream(dev)
            for i, (sim, st) in enumerate(zip(self.members, self.state)):
                s = streams[i] if concurrent else cap
                s.wait_stream(cap)
                with torch.cuda.stream(s):
                    self.advance(sim, st)
                    self._write_back(sim, st, statics[i], rebind=False)
                if concurrent:
                    cap.wait_stream(s)
        # point every member back at its state tensors
        for (sim, st), (names, svars, sst, s64) in zip(zip(self.members, self.state), statics):
            for n in names:
                sim.variables[n].data = svars[n]
            for U, t in s64.items():
                sim._state64[U] = t
                sim._mark_state(U)
            st.clear()
            st.update(sst)
        self.graph = g
        self._streams = streams
        self._statics = statics
        return g

    def step(self, nsteps=1):
        """nsteps replays of the captured step (capture() first)."""
        if self.graph is None:
            raise RuntimeError("capture() the ensemble step before replaying it")
        for _ in range(nsteps):
            self.graph.replay()
            self.steps += 1

    def npoints(self):
        return sum(m.npts for m in self.members)

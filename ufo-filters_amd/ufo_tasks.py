"""Python face of the C task layer (include/ufo_tasks.h), shaped like the
gi.repository.Ufo API the reference's tests drive (tests/test_fft.py:57-112):

    pm = PluginManager()
    fft = pm.get_task("fft"); fft.props.size_x = 1024
    graph = TaskGraph(); graph.connect_nodes(mem_in, fft); ...
    Scheduler(ctx).run(graph)

Scheduler runs a linear graph item by item through ufomi_task_get_requisition /
process / generate, as ufo-core does; Pipeline runs the C batch runner
(ufomi_pipeline_*) that fuses fft ! filter ! ifft.  Both go through libufomi.so
on the GPU; there is no CPU path.  The file tasks `read` (a generator) and
`write` (a sink) are CPU-mode tasks: they also run without a device through
Task.generate_host / Task.process_host.
"""
import ctypes

import numpy as np

import ufomi

LAYOUT_REAL, LAYOUT_COMPLEX = 0, 1
MODE_PROCESSOR, MODE_REDUCTOR, MODE_GENERATOR, MODE_SINK, MODE_GPU, MODE_CPU = 1, 2, 4, 8, 16, 32


class Requisition(ctypes.Structure):
    _fields_ = [("n_dims", ctypes.c_uint), ("dims", ctypes.c_size_t * 3)]


class Buffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("req", Requisition), ("layout", ctypes.c_int), ("dtype", ctypes.c_int),
                ("owned", ctypes.c_int)]


class Error(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("status", ctypes.c_int), ("message", ctypes.c_char * 512)]


class TaskError(RuntimeError):
    """GError of domain UFO_TASK_ERROR: .code (SETUP=0, GET_REQUISITION=1, PROCESS=2) and .status."""

    def __init__(self, err):
        super().__init__(err.message.decode())
        self.code = err.code
        self.status = err.status


_bound = False


def _lib():
    global _bound
    L = ufomi.load()
    if not _bound:
        vp, ci, cu, cs = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_char_p
        B, R, E = ctypes.POINTER(Buffer), ctypes.POINTER(Requisition), ctypes.POINTER(Error)
        sig = {
            "ufomi_task_new": ([cs], vp), "ufomi_task_free": ([vp], None), "ufomi_task_name": ([vp], cs),
            "ufomi_task_set_property": ([vp, cs, cs], ci),
            "ufomi_task_get_property": ([vp, cs, ctypes.c_char_p, ctypes.c_size_t], ci),
            "ufomi_task_list_properties": ([vp], ctypes.POINTER(cs)),
            "ufomi_task_setup": ([vp, vp, E], ci),
            "ufomi_task_get_requisition": ([vp, ctypes.POINTER(B), R, E], ci),
            "ufomi_task_get_num_inputs": ([vp], cu), "ufomi_task_get_num_dimensions": ([vp, cu], cu),
            "ufomi_task_get_mode": ([vp], cu),
            "ufomi_task_process": ([vp, ctypes.POINTER(B), B, R], ci),
            "ufomi_task_generate": ([vp, B, R], ci),
            "ufomi_task_generate_host": ([vp, vp, R], ci), "ufomi_task_process_host": ([vp, vp, R], ci),
            "ufomi_task_output_layout": ([vp], ci), "ufomi_task_last_error": ([vp], cs),
            "ufomi_pipeline_new": ([cs, E], vp), "ufomi_pipeline_free": ([vp], None),
            "ufomi_pipeline_num_tasks": ([vp], ci), "ufomi_pipeline_task": ([vp, ci], vp),
            "ufomi_pipeline_run": ([vp, vp, vp, R, ci, vp, ctypes.c_size_t, ctypes.POINTER(ci), R, E], ci),
            "ufomi_pipeline_used_fused_filter": ([vp], ci),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _bound = True
    return L


TASK_EXPORTS = [
    "ufomi_task_new", "ufomi_task_free", "ufomi_task_name", "ufomi_task_set_property", "ufomi_task_get_property",
    "ufomi_task_list_properties", "ufomi_task_setup", "ufomi_task_get_requisition", "ufomi_task_get_num_inputs",
    "ufomi_task_get_num_dimensions", "ufomi_task_get_mode", "ufomi_task_process", "ufomi_task_generate",
    "ufomi_task_generate_host", "ufomi_task_process_host",
    "ufomi_task_output_layout", "ufomi_task_output_dtype", "ufomi_task_last_error", "ufomi_buffer_new",
    "ufomi_buffer_release", "ufomi_requisition_elements", "ufomi_pipeline_new", "ufomi_pipeline_free",
    "ufomi_pipeline_num_tasks", "ufomi_pipeline_task", "ufomi_pipeline_run", "ufomi_pipeline_used_fused_filter",
    "ufomi_device_alloc", "ufomi_device_free", "ufomi_memcpy", "ufomi_device_memory",
]


# properties of GType gchararray: returned as text, whatever they look like
_STRING_PROPS = {("read", "path"), ("write", "filename")}


class _Props:
    """Attribute access to task properties; names use '_' for '-' like PyGObject."""

    def __init__(self, task):
        object.__setattr__(self, "_task", task)

    def __setattr__(self, name, value):
        self._task.set_property(name.replace("_", "-"), value)

    def __getattr__(self, name):
        return self._task.get_property(name.replace("_", "-"))


def _text(value):
    if isinstance(value, bool):
        return "true" if value else "false"
    if isinstance(value, (list, tuple, np.ndarray)):
        return ",".join(repr(float(v)) for v in value)
    return str(value)


class Task:
    def __init__(self, name):
        self.L = _lib()
        self.h = self.L.ufomi_task_new(name.encode())
        if not self.h:
            raise ValueError("no task named %r" % name)
        self.name = name
        self.owned = True
        self.props = _Props(self)

    def __del__(self):
        try:
            if self.h and getattr(self, "owned", False):
                self.L.ufomi_task_free(self.h)
        except Exception:
            pass

    def set_property(self, name, value):
        rc = self.L.ufomi_task_set_property(self.h, name.encode(), _text(value).encode())
        if rc:
            raise ufomi.UfomiError(rc, self.L.ufomi_last_error().decode())

    def get_property(self, name):
        buf = ctypes.create_string_buffer(1 << 16)
        rc = self.L.ufomi_task_get_property(self.h, name.encode(), buf, len(buf))
        if rc:
            raise ufomi.UfomiError(rc, self.L.ufomi_last_error().decode())
        s = buf.value.decode()
        if (self.L.ufomi_task_name(self.h).decode(), name.replace("_", "-")) in _STRING_PROPS:
            return s
        if s in ("true", "false"):
            return s == "true"
        try:
            return int(s)
        except ValueError:
            pass
        try:
            return float(s)
        except ValueError:
            pass
        if "," in s or s == "":
            return [float(v) for v in s.split(",")] if s else []
        return s

    def list_properties(self):
        arr = self.L.ufomi_task_list_properties(self.h)
        out, i = [], 0
        while arr[i]:
            out.append(arr[i].decode())
            i += 1
        return out

    @property
    def mode(self):
        return self.L.ufomi_task_get_mode(self.h)

    def setup(self, ctx=None):
        """ctx None is allowed for CPU-mode tasks (read, write) on host memory."""
        err = Error()
        if self.L.ufomi_task_setup(self.h, ctx.h if ctx is not None else None, ctypes.byref(err)):
            raise TaskError(err)

    @staticmethod
    def _inputs(buf):
        bufs = buf if isinstance(buf, (list, tuple)) else [buf]
        return (ctypes.POINTER(Buffer) * len(bufs))(*[ctypes.pointer(b) for b in bufs])

    def get_requisition(self, buf=None):
        """buf: a Buffer, a list of them for multi-input tasks, or None for generators; an
        exhausted generator returns n_dims 0."""
        err, req = Error(), Requisition()
        arr = self._inputs(buf) if buf is not None else None
        if self.L.ufomi_task_get_requisition(self.h, arr, ctypes.byref(req), ctypes.byref(err)):
            raise TaskError(err)
        return req

    def generate_host(self, req):
        """One generated item as a float32 numpy array, None at the end; raises on failure."""
        out = np.empty(_req_shape(req), dtype=np.float32)
        if self.L.ufomi_task_generate_host(self.h, out.ctypes.data, ctypes.byref(req)):
            return out
        msg = self.last_error()
        if msg:
            raise RuntimeError("%s.generate failed: %s" % (self.name, msg))
        return None

    def process_host(self, array):
        """Hands one float32 item (numpy [h][w] or [d][h][w]) to a CPU-mode sink."""
        a = np.ascontiguousarray(array, dtype=np.float32)
        req = Requisition()
        req.n_dims = a.ndim
        for i, n in enumerate(reversed(a.shape)):
            req.dims[i] = n
        if not self.L.ufomi_task_process_host(self.h, a.ctypes.data, ctypes.byref(req)):
            raise RuntimeError("%s.process failed: %s" % (self.name, self.last_error()))

    def items(self):
        """Every item of a set-up generator, as numpy arrays."""
        while True:
            req = self.get_requisition()
            if req.n_dims == 0:
                return
            item = self.generate_host(req)
            if item is None:
                return
            yield item

    def process(self, buf, out, req):
        """True when the item was taken.  A reductor may answer False without an error: it takes no
        more input (ufo-core stops feeding it); any other False raises."""
        arr = self._inputs(buf)
        ok = self.L.ufomi_task_process(self.h, arr, ctypes.byref(out), ctypes.byref(req))
        if not ok:
            msg = self.L.ufomi_task_last_error(self.h).decode()
            if msg or not (self.mode & MODE_REDUCTOR):
                raise RuntimeError("%s.process failed: %s" % (self.name, msg))
        return bool(ok)

    def generate(self, out, req):
        return bool(self.L.ufomi_task_generate(self.h, ctypes.byref(out), ctypes.byref(req)))

    def last_error(self):
        return self.L.ufomi_task_last_error(self.h).decode()


class MemoryIn:
    """memory-in: numpy items [n][h][w] float32 (src/ufo-memory-in-task.c)."""

    def __init__(self, array):
        self.array = np.ascontiguousarray(array, dtype=np.float32)
        if self.array.ndim == 2:
            self.array = self.array[None]


class MemoryOut:
    """memory-out: collects every output item as a numpy array."""

    def __init__(self):
        self.items = []

    @property
    def array(self):
        return np.stack(self.items) if self.items else np.zeros((0,))


class PluginManager:
    def get_task(self, name):
        if name == "memory-out":
            return MemoryOut()
        return Task(name)


class TaskGraph:
    def __init__(self):
        self.edges = []

    def connect_nodes(self, a, b):
        self.edges.append((a, b))

    def chain(self):
        src = [a for a, _ in self.edges if all(b is not a for _, b in self.edges)]
        if len(src) != 1:
            raise ValueError("only linear graphs are supported")
        order, node = [src[0]], src[0]
        nxt = {id(a): b for a, b in self.edges}
        while id(node) in nxt:
            node = nxt[id(node)]
            order.append(node)
        return order


_TORCH_DT = {0: "float32", 1: "float16", 2: "bfloat16"}
_DT_CODE = {v: k for k, v in _TORCH_DT.items()}


def _device_tensor(tensor, device, what, dtypes=("float32",)):
    """The checks every tensor passes before its data pointer goes to the C layer: on the
    context's device, one of `dtypes`, contiguous.  Returns the dtype code of the buffer."""
    import torch
    if not isinstance(tensor, torch.Tensor):
        raise TypeError("%s must be a torch tensor, not %s" % (what, type(tensor).__name__))
    if not tensor.is_cuda or tensor.device.index != device:
        raise ValueError("%s must be on cuda:%d (it is on %s)" % (what, device, tensor.device))
    name = str(tensor.dtype).replace("torch.", "")
    if name not in dtypes:
        raise ValueError("%s must be %s (it is %s)" % (what, " or ".join(dtypes), name))
    if not tensor.is_contiguous():
        raise ValueError("%s must be contiguous" % what)
    return _DT_CODE[name]


def _req_shape(req):
    return tuple(int(req.dims[i]) for i in range(req.n_dims))[::-1]


class Scheduler:
    """Runs a linear graph item by item on one context (ufo-core's model)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def run(self, graph):
        import torch
        nodes = graph.chain()
        src, sink, tasks = nodes[0], nodes[-1], nodes[1:-1]
        for t in tasks:
            t.setup(self.ctx)
        if isinstance(src, Task):
            src.setup(self.ctx)
        if isinstance(sink, Task):
            sink.setup(self.ctx)
        keep = []

        def make_buf(tensor, layout):
            b = Buffer()
            b.dtype = _device_tensor(tensor, self.ctx.device, "a task's buffer", tuple(_TORCH_DT.values()))
            b.data = tensor.data_ptr()
            b.req.n_dims = tensor.dim()
            for i, s in enumerate(reversed(tensor.shape)):
                b.req.dims[i] = s
            b.layout = layout
            return b

        def push(stage, tensor, layout):
            """Send one item into tasks[stage:], returns nothing; outputs go to the sink."""
            if stage == len(tasks):
                torch.cuda.synchronize()
                if isinstance(sink, Task):
                    sink.process_host(tensor.cpu().numpy())
                else:
                    sink.items.append(tensor.cpu().numpy())
                return
            t = tasks[stage]
            buf = make_buf(tensor, layout)
            req = t.get_requisition(buf)
            out = torch.empty(_req_shape(req), dtype=torch.float32, device=tensor.device)
            obuf = make_buf(out, LAYOUT_REAL)
            keep.append((tensor, out))
            t.process(buf, obuf, req)
            if not (t.mode & MODE_REDUCTOR):
                push(stage + 1, out, obuf.layout)

        def flush(stage):
            for k in range(stage, len(tasks)):
                t = tasks[k]
                if t.mode & MODE_REDUCTOR:
                    last = reducer_in[k]
                    req = t.get_requisition(last)
                    while True:
                        out = torch.empty(_req_shape(req), dtype=torch.float32, device="cuda")
                        obuf = make_buf(out, LAYOUT_REAL)
                        if not t.generate(obuf, req):
                            break
                        push(k + 1, out, LAYOUT_REAL)

        reducer_in = {}
        dev = torch.device("cuda:%d" % self.ctx.device)
        for item in (src.items() if isinstance(src, Task) else src.array):
            x = torch.from_numpy(np.ascontiguousarray(item, dtype=np.float32)).to(dev)
            buf = make_buf(x, LAYOUT_REAL)
            for k, t in enumerate(tasks):
                if t.mode & MODE_REDUCTOR:
                    reducer_in[k] = buf
            push(0, x, LAYOUT_REAL)
        flush(0)
        torch.cuda.synchronize()


class Pipeline:
    """The C batch runner: `ufo-launch`-style description over device items."""

    def __init__(self, description):
        self.L = _lib()
        err = Error()
        self.h = self.L.ufomi_pipeline_new(description.encode(), ctypes.byref(err))
        if not self.h:
            raise TaskError(err)

    def __del__(self):
        try:
            if self.h:
                self.L.ufomi_pipeline_free(self.h)
        except Exception:
            pass

    def task(self, i):
        t = Task.__new__(Task)
        t.L, t.h, t.name, t.owned = self.L, self.L.ufomi_pipeline_task(self.h, i), None, False
        t.props = _Props(t)
        t.pipeline = self  # the pipeline owns the task; keep it alive
        return t

    @property
    def used_fused_filter(self):
        return bool(self.L.ufomi_pipeline_used_fused_filter(self.h))

    def run(self, ctx, items=None, out_bytes=None):
        """items: torch tensor [n][h][w] float32, contiguous, on the context's device (None
        when the first stage is `read`); returns (tensor of bytes or None when the last stage
        is `write`, n_out, shape).  out_bytes sizes the output buffer (default: 8x the input,
        at most half the free device memory)."""
        import torch
        req = Requisition()
        n, ptr, reqp = 0, None, None
        if items is not None:
            _device_tensor(items, ctx.device, "items")
            if items.dim() != 3:
                raise ValueError("items must be [n][h][w] (it has %d dimensions)" % items.dim())
            n, h, w = items.shape
            req.n_dims = 2
            req.dims[0], req.dims[1], req.dims[2] = w, h, 1
            ptr, reqp = items.data_ptr(), ctypes.byref(req)
        n_tasks = self.L.ufomi_pipeline_num_tasks(self.h)
        last = self.L.ufomi_task_name(self.L.ufomi_pipeline_task(self.h, n_tasks - 1)).decode()
        out, cap = None, 0
        if last != "write":
            # without out_bytes: room for 8x the input, but at most half the free device
            # memory (the stages' own batches need the rest)
            cap = out_bytes
            if not cap:
                free, _ = torch.cuda.mem_get_info(ctx.device)
                cap = min(items.numel() * 4 * 8 if items is not None else 1 << 30, free // 2)
            out = torch.empty(cap, dtype=torch.uint8, device="cuda:%d" % ctx.device)
        n_out, oreq, err = ctypes.c_int(), Requisition(), Error()
        rc = self.L.ufomi_pipeline_run(self.h, ctx.h, ptr, reqp, n, out.data_ptr() if out is not None else None, cap,
                                       ctypes.byref(n_out), ctypes.byref(oreq), ctypes.byref(err))
        if rc:
            raise TaskError(err)
        return out, n_out.value, _req_shape(oreq)

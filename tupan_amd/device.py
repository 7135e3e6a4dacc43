"""Device-resident entry points over torch tensors (ROCm).

PyTorch is used only as an allocator and stream source: tensors must be
contiguous float64 on a HIP device; their data pointers go straight to the
``tupan_<name>_dev`` C ABI, which launches on torch's current stream (or the
one given).  Nothing is copied.
"""
import torch

from . import _lib, spec


def _field(obj, name):
    return obj[name] if isinstance(obj, dict) else getattr(obj, name)


def _check_tensor(t, n, what):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % what)
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("%s must be a contiguous float64 device tensor" % what)
    if t.dim() != 1 or (n is not None and t.numel() != n):
        raise ValueError("%s has the wrong length" % what)


def current_stream_ptr(device=None):
    return torch.cuda.current_stream(device).cuda_stream


def compute(name, ips, jps, *args, mode="fp64", out=None, stream=None,
            status=None):
    """Evaluate kernel ``name``; returns the tuple of output tensors.

    ``status``: optional int32 device tensor of length 1, OR-ed on the
    device (no synchronisation) with 1 when a Sakura Kepler solve hit its
    cap and with 2 when an output holds NaN or inf
    (tupan_check_finite_dev).
    """
    k = spec.KERNELS[name]
    it = [_field(ips, f) for f in k.fields]
    jt = [_field(jps, f) for f in k.fields]
    ni = it[0].numel()
    nj = jt[0].numel()
    for f, t in zip(k.fields, it):
        _check_tensor(t, ni, "i." + f)
    for f, t in zip(k.fields, jt):
        _check_tensor(t, nj, "j." + f)
    dev = it[0].device
    if out is None:
        out = tuple(torch.empty(ni, dtype=torch.float64, device=dev)
                    for _ in k.outs)
    for o in out:
        _check_tensor(o, ni, "output")
    st = stream if stream is not None else current_stream_ptr(dev)
    sptr = None
    if status is not None:
        if status.dtype != torch.int32 or not status.is_cuda:
            raise ValueError("status must be an int32 device tensor")
        sptr = status.data_ptr()
    _lib.call_dev(name, mode, ni, [t.data_ptr() for t in it], nj,
                  [t.data_ptr() for t in jt], [o.data_ptr() for o in out],
                  args, stream=st, status_ptr=sptr)
    if sptr is not None:
        _lib.check_finite_dev(ni, [o.data_ptr() for o in out], sptr, st)
    return out


def to_device(ps, fields, device="cuda"):
    """Copy the named SoA fields of a host particle set to device tensors."""
    return {f: torch.as_tensor(_field(ps, f), dtype=torch.float64,
                               device=device).contiguous() for f in fields}

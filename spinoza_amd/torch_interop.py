"""Zero-copy torch views of a state's device arrays (torch is plumbing only)."""
import ctypes


def device_views(state):
    """(reals, imags) torch tensors aliasing the state's HBM arrays."""
    import torch
    re_ptr, im_ptr = state.device_ptrs()
    n = len(state)
    dtype = torch.float64 if state.precision == "f64" else torch.float32

    class _Holder:
        def __init__(self, ptr):
            esz = 8 if dtype == torch.float64 else 4
            self.__cuda_array_interface__ = {
                "shape": (n,), "typestr": "<f8" if esz == 8 else "<f4",
                "data": (ptr, False), "version": 3, "strides": None}

    return (torch.as_tensor(_Holder(re_ptr), device="cuda"),
            torch.as_tensor(_Holder(im_ptr), device="cuda"))

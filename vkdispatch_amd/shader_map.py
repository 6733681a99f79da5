"""``vd.map``: mapping functions for the FFT read / write / kernel phases.

Same contract as the reference (vkdispatch/shader/map.py:10-69): the wrapped
function's parameters are annotated ``vc.Buffer[T]`` / ``vc.Var[T]`` /
``vc.Const[T]``; those become extra kernel arguments; equality and hashing are
per instance (a new ``vd.map`` object means a new kernel).
"""
from __future__ import annotations

import dataclasses
import inspect
import uuid
from typing import Callable, List


@dataclasses.dataclass(frozen=True)
class MappingFunction:
    buffer_types: List
    return_type: object
    mapping_function: Callable
    instance_id: uuid.UUID = dataclasses.field(default_factory=uuid.uuid4, compare=False)

    def __hash__(self) -> int:
        return hash(self.instance_id)

    def __eq__(self, other):
        return isinstance(other, MappingFunction) and self.instance_id == other.instance_id

    def callback(self, *args):
        """Run the user function on traced arguments (inside a C++ scope)."""
        from . import codegen as vc
        with vc.scope_block():
            return self.mapping_function(*args)


def map(func: Callable = None, return_type=None, input_types: List = None):  # noqa: A001
    """Wrap ``func`` as a MappingFunction (usable as ``@vd.map`` or ``vd.map(func)``)."""
    if func is None:
        return lambda f: map(f, return_type=return_type, input_types=input_types)
    sig = inspect.signature(func)
    if return_type is None and sig.return_annotation is not inspect.Parameter.empty:
        return_type = sig.return_annotation
    if input_types is None:
        input_types = []
        hints = None
        for param in sig.parameters.values():
            ann = param.annotation
            if ann is inspect.Parameter.empty:
                raise ValueError("All parameters must be annotated")
            if isinstance(ann, str):  # postponed evaluation (from __future__ import annotations)
                if hints is None:
                    import typing
                    hints = typing.get_type_hints(func)
                ann = hints.get(param.name, ann)
            if not hasattr(ann, "__args__"):
                raise TypeError(f"Argument '{param.name}: {ann}' must have a type annotation")
            input_types.append(ann)
    return MappingFunction(buffer_types=list(input_types), return_type=return_type, mapping_function=func)

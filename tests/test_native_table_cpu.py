"""CPU tests of the reference-side backend module (vkdispatch_amd/native_table.py):
the compile and error paths of stage_compute_plan_create through ctypes, and the
argument checks of record / submit.  Nothing is launched (no GPU here)."""
import pytest

import vkdispatch_amd as vd
from vkdispatch_amd import native_table as nt


@pytest.fixture
def ctx():
    # a context object without a device: enough for the compile path
    nt.clear_error()
    h = nt._new(nt._Context(0, 0))
    yield h
    nt.context_destroy(h)
    nt.clear_error()


def test_error_string_is_none_until_an_error(ctx):
    assert nt.get_error_string() == 0  # the reference returns 0 when there is no error
    assert nt.buffer_create(ctx, 0, 1) == 0
    assert "greater than zero" in nt.get_error_string()


def test_plan_create_with_bad_hip_source_returns_zero_with_log(ctx):
    good = vd.fft.make_fft_shader((8, 256), None).source
    first, rest = good.split("\n", 1)
    bad = first + "\n" + rest.replace("vkx::fft_body", "vkx::no_such_body")
    assert nt.stage_compute_plan_create(ctx, bad.encode(), [0], 0, b"bad") == 0
    msg = nt.get_error_string()
    assert msg.startswith("Failed to compile HIP kernel 'bad'")
    assert "no_such_body" in msg  # the compiler log is part of the message


def test_plan_create_needs_the_launch_line(ctx):
    src = vd.fft.make_fft_shader((8, 256), None).source.split("\n", 1)[1]
    assert nt.stage_compute_plan_create(ctx, src, [0], 0, "k") == 0
    assert "launch description" in nt.get_error_string()


def test_plan_create_compiles_a_planner_kernel_then_needs_a_device(ctx):
    """A translation unit from the planner compiles; on a host without a GPU the module
    load is the step that fails, and it fails with the runtime's message."""
    src = vd.fft.make_fft_shader((4, 1024), None).source
    assert nt.stage_compute_plan_create(ctx, src, [0], 0, "fft1024") == 0
    msg = nt.get_error_string()
    assert msg.startswith("stage_compute_plan_create:") and "hipModuleLoadData" in msg, msg


def test_record_and_submit_reject_bad_handles(ctx):
    nt.stage_compute_record(12345, 67890, 0, 1, 1, 1)
    assert "Invalid command list or compute plan" in nt.get_error_string()
    nt.clear_error()
    assert nt.command_list_submit(424242, b"", 1, 0) is True
    assert "Invalid command list" in nt.get_error_string()
    assert nt.descriptor_set_create(99999) == 0
    assert nt.command_list_create(99999) == 0

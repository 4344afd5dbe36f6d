"""Test helper: build plans under a given schedule of the engine."""
import contextlib
import dataclasses


@contextlib.contextmanager
def schedule(engine, **kw):
    prev = engine.schedule
    engine.set_schedule(dataclasses.replace(prev, **kw))
    try:
        yield
    finally:
        engine.set_schedule(prev)

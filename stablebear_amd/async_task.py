"""Task polling with an optional progress bar (``stablebear/async_task.py``).

Polls every 5 ms rather than 50 ms so small jobs return promptly; errors a
job stored surface from ``wait_for``.
"""

from __future__ import annotations

_POLL_MS = 5


def _wait_for_task(task, verbose=True):
    progress = None
    if verbose:
        from tqdm import tqdm

        progress = tqdm(total=task.work_total(), unit_scale=True, unit=task.work_step_unit(),
                        desc=task.work_step_desc())
    while not task.wait_for(_POLL_MS):
        if progress is not None:
            progress.update(task.work_completed() - progress.n)
    if progress is not None:
        progress.update(task.work_completed() - progress.n)
        progress.close()


def _run_task(task_fn, verbose=True):
    task = None
    try:
        task = task_fn()
        _wait_for_task(task, verbose=verbose)
    finally:
        if task is not None:
            task.request_stop()
            try:
                _wait_for_task(task, verbose=False)
            except BaseException:
                pass

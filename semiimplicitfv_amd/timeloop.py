"""Host time loop mirroring runTimeLoop (src/RKTimeStepping.cpp:241-356).

stepFn(targetDt) -> dt is called with targetDt clamped to the next output
time and the end time, exactly as the reference loop does; on_output(time)
runs where the reference writes VTK output.  VTK, printing and the NaN
abort are the caller's business.
"""
from __future__ import annotations


def run_time_loop(config, stepFn, endTime: float, outputInterval: float, on_output=None, max_steps=None):
    time = 0.0
    nextOutput = outputInterval
    steps = 0
    while time < endTime:
        targetDt = endTime - time
        if nextOutput < endTime:
            targetDt = min(targetDt, nextOutput - time)
        config.time = time
        dt = stepFn(targetDt)
        time += dt
        config.step += 1
        steps += 1
        if time >= nextOutput - 1e-12 * outputInterval:
            if on_output:
                on_output(time)
            nextOutput += outputInterval
        if max_steps is not None and steps >= max_steps:
            break
    return time, steps

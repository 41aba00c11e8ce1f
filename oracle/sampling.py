"""CPU restatement of the engine's trajectory sampling schedule.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Reference behaviour:
  OutputManager::initialize sets next_write_time_ = write_interval_dt_
    (src/core/integrator/OutputManager.cpp:161);
  SimulationEngine::run logs after a step when current_step_ == 0 or
    should_write(new_time) (src/core/integrator/SimulationEngine.cpp:270-281);
  should_write(t) is t >= next_write_time_ (OutputManager.cpp:246-253);
  log_step(t) sets next_write_time_ = t + write_interval_dt_ (OutputManager.cpp:230);
  new_time is current_time_ + dt for the fixed-step integrators, and current_time_
    takes new_time (SimulationEngine.cpp:283).
"""


class OutputSchedule:
    """OutputManager's time-based sampling state."""

    def __init__(self, write_interval):
        self.write_interval = write_interval
        self.next_write_time = write_interval   # OutputManager.cpp:161

    def should_write(self, t):
        return t >= self.next_write_time        # OutputManager.cpp:252

    def log_step(self, t):
        self.next_write_time = t + self.write_interval   # OutputManager.cpp:230


def snapshot_steps(t0, dt, n_steps, write_interval, step_offset=0, first=True, next_write=None):
    """The engine steps after which a snapshot is taken, over n_steps fixed steps from t0.

    Returns ([(step, t), ...], next_write_time after the run)."""
    sched = OutputSchedule(write_interval)
    if next_write is not None:
        sched.next_write_time = next_write
    out = []
    t = t0
    for k in range(n_steps):
        new_time = t + dt
        if (k == 0 and first) or sched.should_write(new_time):
            sched.log_step(new_time)
            out.append((step_offset + k, new_time))
        t = new_time
    return out, sched.next_write_time

"""Job parameters (altar/altar/simulations/Job.py:24-46) and the random stream seed."""
from dataclasses import dataclass


@dataclass
class Job:
    chains: int = 1024      # number of Markov chains
    steps: int = 20         # Metropolis steps per beta stage
    tolerance: float = 1e-3  # stop when beta + tolerance >= 1
    seed: int = 0            # Philox key of every random stream
    gpus: int = 1

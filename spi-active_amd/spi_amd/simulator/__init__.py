"""HIP simulator backend behind the BaseSimulator surface of spigym/simulator."""

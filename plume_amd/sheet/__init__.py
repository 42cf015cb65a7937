"""Sheet frames from raw frames (sheet/ of the reference): the union content box and its square crop,
bicubic resampling to the frame size, the soften blur, the shade layout's thin-pixel easing and divisor, and
the 8-bit quantization. Host-side numpy and Pillow on a run's few small frames, after the ray march has
brought each raw frame to the host."""
